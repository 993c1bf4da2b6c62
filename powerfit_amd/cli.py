"""``powerfit``-compatible command line: fit an atomic model into a cryo-EM map
with the exhaustive 6-D LCC search on an MI355X.

Follows the reference's ``main`` (``src/powerfit_em/powerfit.py:25-319``):
the same positional arguments and options, the same map pipeline
(resample when the map is finer than ``resampling-rate`` x Nyquist, trim at
10 % of the maximum, extend every axis to the next 2,3,5,7-smooth length),
template and mask from the model on the map's grid, optional core weighting,
the search, the analysis and the same output files in ``--directory``:
``powerfit.log``, ``lcc.mrc``, ``solutions.out`` and ``fit_<n>.pdb``.

The search always runs on the GPU (HIP): ``-g/--gpu`` is accepted for
compatibility and ``-p/--nproc`` is ignored; ``--device`` picks the GPU.

    python -m powerfit_amd.cli target.map 8.0 model.pdb -a 10 -l -cw -d out/
"""
import logging
from argparse import ArgumentParser, FileType
from os.path import abspath, join
from sys import argv, stdout
from time import time

from .analyzer import Analyzer
from .helpers import determine_core_indices, fisher_sigma, mkdir_p, write_fits_to_pdb
from .powerfitter import PowerFitter
from .rotations import proportional_orientations, quat_to_rotmat
from .structure import Structure
from .volume import Volume, extend, nearest_multiple2357, resample, structure_to_shape_like, trim


def parse_args(args=None):
    p = ArgumentParser(description="Rigid-body fitting of an atomic model into a cryo-EM map "
                                   "(exhaustive LCC search on the GPU).")
    p.add_argument("target", type=FileType("rb"),
                   help="Target density map to fit the model in (CCP4 or MRC).")
    p.add_argument("resolution", type=float, help="Resolution of map in angstrom")
    p.add_argument("template", type=FileType("r"),
                   help="Atomic model to be fitted in the density (PDB or mmCIF).")
    p.add_argument("-a", "--angle", dest="angle", type=float, default=10, metavar="<float>",
                   help="Rotational sampling density in degree.")
    p.add_argument("-l", "--laplace", dest="laplace", action="store_true",
                   help="Use the Laplace pre-filter density data.")
    p.add_argument("-cw", "--core-weighted", dest="core_weighted", action="store_true",
                   help="Use core-weighted local cross-correlation score.")
    p.add_argument("-nr", "--no-resampling", dest="no_resampling", action="store_true",
                   help="Do not resample the density map.")
    p.add_argument("-rr", "--resampling-rate", dest="resampling_rate", type=float, default=2,
                   metavar="<float>", help="Resampling rate compared to Nyquist.")
    p.add_argument("-nt", "--no-trimming", dest="no_trimming", action="store_true",
                   help="Do not trim the density map.")
    p.add_argument("-tc", "--trimming-cutoff", dest="trimming_cutoff", type=float, default=None,
                   metavar="<float>", help="Intensity cutoff to which the map will be trimmed. "
                   "Default is 10 percent of maximum intensity.")
    p.add_argument("-c", "--chain", dest="chain", type=str, default=None, metavar="<char>",
                   help="The chain IDs of the structure to be fitted, comma separated.")
    p.add_argument("-d", "--directory", dest="directory", type=abspath, default=".",
                   metavar="<dir>", help="Directory where the results are stored.")
    p.add_argument("-n", "--num", dest="num", type=int, default=10, metavar="<int>",
                   help="Number of models written to file.")
    p.add_argument("-g", "--gpu", dest="gpu", action="store_true",
                   help="Accepted for compatibility: the search always runs on the GPU.")
    p.add_argument("-p", "--nproc", dest="nproc", type=int, default=1, metavar="<int>",
                   help="Accepted for compatibility and ignored.")
    p.add_argument("--device", dest="device", type=int, default=0, metavar="<int>",
                   help="GPU ordinal of the search.")
    return p.parse_args(args)


def write(line):
    """Line to stdout (when a terminal) and to the log file."""
    if stdout.isatty():
        print(line)
    logging.info(line)


def prepare_target(target, resolution, args):
    """The map pipeline of ``powerfit.py:216-235``: resample, trim, extend."""
    if not args.no_resampling:
        factor = 2 * args.resampling_rate * target.voxelspacing / resolution
        if factor < 0.9:
            target = resample(target, factor)
            write(("Shape after resampling:" + " {:d}" * 3).format(*target.shape))
    if not args.no_trimming:
        if args.trimming_cutoff is None:
            args.trimming_cutoff = target.array.max() / 10
        target = trim(target, args.trimming_cutoff)
        write(("Shape after trimming:" + " {:d}" * 3).format(*target.shape))
    target = extend(target, [nearest_multiple2357(n) for n in target.shape])
    write(("Shape after extending:" + " {:d}" * 3).format(*target.shape))
    return target


def main(args=None):
    time0 = time()
    args = parse_args(args)
    mkdir_p(args.directory)
    logging.basicConfig(filename=join(args.directory, "powerfit.log"), level=logging.INFO,
                        format="%(asctime)s %(message)s", force=True)
    logging.info(" ".join(argv))

    write("Target file read from: {:s}".format(abspath(args.target.name)))
    target = Volume.fromfile(args.target.name)
    args.target.close()
    write("Target resolution: {:.2f}".format(args.resolution))
    resolution = args.resolution
    write(("Initial shape of density:" + " {:d}" * 3).format(*target.shape))
    target = prepare_target(target, resolution, args)

    write("Template file read from: {:s}".format(abspath(args.template.name)))
    structure = Structure.fromfile(args.template.name)
    args.template.close()
    if args.chain is not None:
        write("Selecting chains: " + args.chain)
        structure = structure.select("chain", args.chain.split(","))
    if structure.data.size == 0:
        raise ValueError("No atoms were selected.")

    # the model's centre onto the map's origin, then template and mask on its grid
    structure.translate(target.origin - structure.coor.mean(axis=1))
    template = structure_to_shape_like(target, structure.coor, resolution=resolution,
                                       weights=structure.atomnumber, shape="vol",
                                       device=args.device)
    mask = structure_to_shape_like(target, structure.coor, resolution=resolution, shape="mask",
                                   device=args.device)

    write("Reading in rotations.")
    q, w, degree = proportional_orientations(args.angle)
    rotmat = quat_to_rotmat(q)
    write("Requested rotational sampling density: {:.2f}".format(args.angle))
    write("Real rotational sampling density: {:}".format(degree))

    if args.core_weighted:
        write("Calculating core-weighted mask.")
        mask.array = determine_core_indices(mask.array)

    pf = PowerFitter(target, laplace=args.laplace)
    pf._rotations = rotmat
    pf._template = template
    pf._mask = mask
    pf._nproc = args.nproc
    pf.directory = args.directory
    pf._queues = [args.device]
    write("Using GPU-accelerated search.")
    write("Starting search")
    time1 = time()
    pf.scan()
    write("Time for search: {:.0f}m {:.0f}s".format(*divmod(time() - time1, 60)))

    write("Analyzing results")
    mv = structure_to_shape_like(target, structure.coor, resolution=resolution,
                                 radii=structure.rvdw, shape="mask",
                                 device=args.device).array.sum() * target.voxelspacing ** 3
    z_sigma = fisher_sigma(mv, resolution)
    analyzer = Analyzer(pf._lcc, rotmat, pf._rot, voxelspacing=target.voxelspacing,
                        origin=target.origin, z_sigma=z_sigma, device=args.device)

    write("Writing solutions to file.")
    Volume(pf._lcc, target.voxelspacing, target.origin).tofile(join(args.directory, "lcc.mrc"))
    analyzer.tofile(join(args.directory, "solutions.out"))

    write("Writing PDBs to file.")
    n = min(args.num, len(analyzer.solutions))
    write_fits_to_pdb(structure, analyzer.solutions[:n], basename=join(args.directory, "fit"))
    write("Total time: {:.0f}m {:.0f}s".format(*divmod(time() - time0, 60)))
    return pf, analyzer


if __name__ == "__main__":
    main()
