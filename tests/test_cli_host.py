"""Host side of the powerfit-compatible command line: options and the map
pipeline before the search (powerfit.py:25-166, 216-235)."""
import numpy as np

from powerfit_amd import cli
from powerfit_amd.volume import Volume, is_multiple2357, write_map


def test_options_match_reference_defaults(tmp_path):
    m = tmp_path / "t.mrc"
    write_map(str(m), Volume(np.ones((4, 4, 4), np.float32), 1.0, (0, 0, 0)))
    p = tmp_path / "m.pdb"
    p.write_text("END\n")
    a = cli.parse_args([str(m), "6.5", str(p)])
    assert (a.angle, a.laplace, a.core_weighted, a.no_resampling, a.resampling_rate,
            a.no_trimming, a.trimming_cutoff, a.chain, a.num, a.nproc) == \
        (10, False, False, False, 2, False, None, None, 10, 1)
    a = cli.parse_args([str(m), "6.5", str(p), "-a", "20", "-l", "-cw", "-nr", "-nt", "-c",
                        "A,B", "-n", "3", "-g", "-p", "4"])
    assert (a.angle, a.laplace, a.core_weighted, a.no_resampling, a.no_trimming, a.chain,
            a.num, a.gpu, a.nproc) == (20, True, True, True, True, "A,B", 3, True, 4)
    a.target.close()
    a.template.close()


def test_map_pipeline_resample_trim_extend():
    rng = np.random.default_rng(3)
    arr = np.zeros((40, 44, 50))
    arr[10:25, 12:30, 8:33] = rng.random((15, 18, 25)) + 0.5
    vol = Volume(arr, 1.0, (2.0, 4.0, 6.0))

    class A:
        no_resampling = False
        resampling_rate = 2
        no_trimming = False
        trimming_cutoff = None
    out = cli.prepare_target(vol, 8.0, A)          # factor 0.5: resampled
    assert out.voxelspacing == 2.0
    assert all(is_multiple2357(n) for n in out.shape)
    assert A.trimming_cutoff > 0
    A.no_resampling = True
    A.trimming_cutoff = None
    out = cli.prepare_target(vol, 8.0, A)
    assert out.voxelspacing == 1.0 and all(is_multiple2357(n) for n in out.shape)
    # trimmed to the slabs above 10 % of the maximum plus 2 voxels, data at the low corner
    assert np.array_equal(out.array[:19, :22, :29], arr[8:27, 10:32, 6:35])
