"""The powerfit-compatible command line (powerfit.py:189-319) end to end on a
synthetic map and model: map pipeline, template and mask on the GPU, search,
analysis, output files; the LCC and rotation maps checked against the FP64
oracle on the same prepared inputs."""
import os

import numpy as np
import pytest

from oracle import correlator as oc
from powerfit_amd import cli
from powerfit_amd.structure import Structure
from powerfit_amd.synth import make_chain, random_rotation, splat
from powerfit_amd.volume import Volume, write_map

from parity import assert_scan_parity

pytestmark = pytest.mark.gpu


def write_pdb_chain(path, xyz, elements):
    with open(path, "w") as fh:
        for k, (p, e) in enumerate(zip(xyz.T, elements), start=1):
            name = {"C": "CA", "N": "N", "O": "O", "S": "SD"}[e]
            fh.write(f"ATOM  {k:>5d}  {name:<3s} ALA {'A' if k % 97 else 'B'}{k // 4 + 1:>4d}    "
                     f"{p[0]:8.3f}{p[1]:8.3f}{p[2]:8.3f}  1.00 20.00          {e:>2s}  \n")
        fh.write("END\n")


def test_cli_end_to_end(tmp_path):
    rng = np.random.default_rng(6100)
    xyz, weights = make_chain(400, 12.0, rng)
    elements = np.array([{6.0: "C", 7.0: "N", 8.0: "O", 16.0: "S"}[w] for w in weights])
    vs, res = 2.0, 8.0
    shape = (30, 32, 34)
    rot = random_rotation(rng)
    center = np.array([34 * vs / 2, 32 * vs / 2, 30 * vs / 2])
    origin = np.array([-5.0, 3.0, 12.0])
    moved = rot @ xyz + center[:, None]
    sigma = res / (np.sqrt(2.0) * np.pi) / vs
    dens = splat(moved / vs, weights, sigma, shape)
    dens += rng.normal(0, 0.02 * dens.max(), shape)
    write_map(str(tmp_path / "target.mrc"), Volume(dens.astype(np.float32), vs, tuple(origin)))
    write_pdb_chain(str(tmp_path / "model.pdb"), xyz + origin[:, None] + 7.0, elements)
    out = tmp_path / "out"
    print("cli: 30x32x34 map, 400-atom model, 20 degrees", flush=True)
    pf, an = cli.main([str(tmp_path / "target.mrc"), str(res), str(tmp_path / "model.pdb"),
                       "-a", "20", "-d", str(out), "-n", "3", "-cw"])
    # as powerfit.py:313-317: min(num, number of solutions) models are written
    nfit = min(3, len(an.solutions))
    assert nfit >= 1
    for f in ("powerfit.log", "lcc.mrc", "solutions.out", "fit_1.pdb", f"fit_{nfit}.pdb"):
        assert (out / f).exists(), f
    assert not (out / f"fit_{nfit + 1}.pdb").exists()
    lcc_file = Volume.fromfile(str(out / "lcc.mrc"))
    assert lcc_file.shape == pf._lcc.shape and np.array_equal(lcc_file.array, pf._lcc)
    assert all(n == cli.nearest_multiple2357(n) for n in pf._lcc.shape)
    # parity of the search on the inputs the CLI prepared
    R = pf._rotations
    ref = oc.parallel_scan(np.asarray(pf._target.array, np.float64), pf._template.array,
                           pf._mask.array, R, laplace=False, nproc=min(8, os.cpu_count() or 1))
    err, dec = assert_scan_parity(pf._lcc, pf._rot, ref)
    print(f"  max err {err:.2e}, decisive voxels {dec}, solutions {len(an.solutions)}", flush=True)
    fit = Structure.fromfile(str(out / "fit_1.pdb"))
    assert fit.coor.shape == (3, 400)
    best = an.solutions[0]
    assert best[0] == pf._lcc.max()
    # the best solution's model is the input rotated about its centre and moved
    model = Structure.fromfile(str(tmp_path / "model.pdb"))
    c = model.coor - model.coor.mean(axis=1, keepdims=True)
    R1 = np.asarray(best[6:], dtype=np.float64).reshape(3, 3)
    assert np.allclose(fit.coor, R1 @ c + np.asarray(best[3:6])[:, None], atol=2e-3)
