// Host build of the FFT stage code in pf_fft.h: runs every Stockham stage
// sequentially on the CPU (each butterfly of a stage is independent, so this
// is exactly what a block computes).  Exposed through a tiny C function so
// CPU tests can compare the device butterflies' arithmetic and indexing with
// numpy.fft without a GPU.
#include <cmath>
#include <vector>

#include "pf_fft.h"

extern "C" int pf_host_fft(int n, int dir, const float* in, float* out) {
    pf::Plan1D p;
    if (!pf::make_plan(n, &p)) return -1;
    std::vector<float2> tw(n), a(n), b(n);
    for (int m = 0; m < n; ++m) {
        const double ang = -2.0 * M_PI * (double)m / (double)n;
        tw[m] = pf::c2((float)std::cos(ang), (float)std::sin(ang));
        a[m] = pf::c2(in[2 * m], in[2 * m + 1]);
    }
    float2* src = a.data();
    float2* dst = b.data();
    int Ns = 1;
    for (int s = 0; s < p.nstages; ++s) {
        const int R = p.radix[s];
        for (int j = 0; j < n / R; ++j) {
            if (dir < 0) pf::stage_dispatch<-1>(src, dst, R, j, Ns, n, 1, tw.data());
            else pf::stage_dispatch<1>(src, dst, R, j, Ns, n, 1, tw.data());
        }
        float2* t = src;
        src = dst;
        dst = t;
        Ns *= R;
    }
    for (int m = 0; m < n; ++m) {
        out[2 * m] = src[m].x;
        out[2 * m + 1] = src[m].y;
    }
    return p.nstages;
}

extern "C" int pf_host_plan(int n, int* radix, int max) {
    pf::Plan1D p;
    if (!pf::make_plan(n, &p)) return -1;
    for (int s = 0; s < p.nstages && s < max; ++s) radix[s] = p.radix[s];
    return p.nstages;
}
