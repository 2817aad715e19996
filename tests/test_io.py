"""Reference run files: datafile parsing, per-view raw projections with the
x/y flip, angle files, raw volumes, and an end-to-end CLI run (GPU) against
the oracle on the same files."""
import os

import numpy as np
import pytest

from newrecon_amd.datafile import (DatafileError, load_angles, load_datafile, load_projections,
                                   output_file, read_param, save_projections, view_file)
from newrecon_amd.geometry import Geometry

DATAFILE = """# OSTR run (keys of Tomo_OSTR.cu:2578-2650)
image_size_x 200  image_size_y 100  image_size_z 20
Voxel_size_x_mm 0.5  Voxel_size_y_mm 0.5  Voxel_size_z_mm 1
object_xpos_mm -50  object_ypos_mm -50  object_zpos_mm 0
detector_x0_mm -71.68  detector_y0_mm -71.68
detector_COR_mm 0  source_COR_mm 660  source_Y_shift_mm 0
angular_samples 9
BACKPRJ_ThreX 200 BACKPRJ_ThreY 1 BACKPRJ_GridX 100 BACKPRJ_GridY 20 nBatchXdim 1
bin_samples_x 256  bin_samples_y 128
Bin_size_x_mm 0.56  Bin_size_y_mm 0.56
nBatchBINSx 1 nBatchBINSy 1 PRJ_ThreX 16 PRJ_ThreY 16 PRJ_GridX 16 PRJ_GridY 8
Readout_Every_X_Iters 2
Recon_Method 1
Beta 1000  Delta 0.0005
iteration_number 4  subset_number 3
ini_guess {d}/none.raw
angle_name {d}/angles.ini
projection_folder {d}/prj/  projection_name proj.
scatter_folder {d}/scat/  scatter_name scat.
blank_folder {d}/blank/  blank_name blank.
output_recon_folder {d}/out/  output_recon_name recon_
"""


def write_run(tmp, y, b, r, angles):
    for sub in ("prj", "scat", "blank", "out"):
        os.makedirs(os.path.join(tmp, sub), exist_ok=True)
    save_projections(os.path.join(tmp, "prj") + "/", "proj.", y)
    save_projections(os.path.join(tmp, "scat") + "/", "scat.", r)
    save_projections(os.path.join(tmp, "blank") + "/", "blank.", b)
    angles.astype(np.float32).tofile(os.path.join(tmp, "angles.ini"))
    path = os.path.join(tmp, "run.dat")
    with open(path, "w") as f:
        f.write(DATAFILE.format(d=tmp))
    return path


def test_datafile_parse(tmp_path):
    path = os.path.join(tmp_path, "run.dat")
    with open(path, "w") as f:
        f.write(DATAFILE.format(d=tmp_path))
    run = load_datafile(path)
    g = run.geom
    assert (g.nx, g.ny, g.nz, g.nbx, g.nby) == (200, 100, 20, 256, 128)
    assert g.x_d0 == float(np.float32(-71.68)) and g.source_r == 660.0
    assert run.n_views == 9 and run.readout_every == 2
    assert run.cfg.prior == "lange" and run.cfg.n_subsets == 3 and run.cfg.n_iters == 4
    assert run.cfg.gamma_mode == "ray_length" and run.cfg.guard is False
    assert run.paths["projection_name"] == "proj."


@pytest.mark.parametrize("edit,msg", [
    (("Beta 1000", "Beta #"), "Beta"),
    (("Delta 0.0005", "Delta abc"), "Delta"),
    (("subset_number 3", "subset_number 2"), "multiple"),
    (("Recon_Method 1", "Recon_Method 2"), "Recon_Method"),
    (("image_size_x 200", ""), "image_size_x"),
])
def test_datafile_errors(tmp_path, edit, msg):
    path = os.path.join(tmp_path, "run.dat")
    with open(path, "w") as f:
        f.write(DATAFILE.format(d=tmp_path).replace(*edit))
    with pytest.raises(DatafileError, match=msg):
        load_datafile(path)
    with pytest.raises(DatafileError):
        load_datafile(os.path.join(tmp_path, "missing.dat"))


def test_datafile_values_parse_like_atoi_atof(tmp_path):
    """readdata.cu converts with atoi/atof (:114, :300): the leading number of
    a value counts and the rest is ignored."""
    from newrecon_amd.datafile import atof, atoi
    assert atoi("512.0") == 512 and atoi("-3x") == -3 and atoi("+7") == 7 and atoi("-") == 0
    assert atof("0.1mm") == 0.1 and atof("1e3,") == 1000.0 and atof(".5") == 0.5
    assert atof("-2.5e-1z") == -0.25 and atof("+") == 0.0
    path = os.path.join(tmp_path, "run.dat")
    text = (DATAFILE.format(d=tmp_path).replace("image_size_x 200", "image_size_x 200.0")
            .replace("Voxel_size_z_mm 1", "Voxel_size_z_mm 1mm")
            .replace("Beta 1000", "Beta 1e3,"))
    with open(path, "w") as f:
        f.write(text)
    run = load_datafile(path)
    assert run.geom.nx == 200 and run.geom.dz == 1.0 and run.cfg.beta == 1000.0


def test_projection_files_flip_and_naming(tmp_path):
    nbx, nby = 5, 3
    prj = np.arange(2 * nbx * nby, dtype=np.float32).reshape(2, nbx, nby)
    folder = str(tmp_path) + "/"
    save_projections(folder, "p.", prj)
    assert os.path.exists(folder + "p.0001") and os.path.exists(folder + "p.0002")
    raw = np.fromfile(folder + "p.0002", dtype=np.float32)
    # the file is [nby][nbx], x fastest: element (i, j) sits at j*nbx + i
    assert raw[2 * nbx + 4] == prj[1, 4, 2]
    np.testing.assert_array_equal(load_projections(folder, "p.", 2, nbx, nby), prj)
    assert view_file("a/", "b.", 11) == "a/b.0012"
    assert output_file("o/", "r_", 3) == "o/r_3"
    os.remove(folder + "p.0002")
    with pytest.raises(DatafileError, match="missing"):
        load_projections(folder, "p.", 2, nbx, nby)
    np.zeros(4, np.float32).tofile(folder + "p.0002")
    with pytest.raises(DatafileError, match="truncated"):
        load_projections(folder, "p.", 2, nbx, nby)


def test_angles_and_raw_volume(tmp_path):
    import torch
    from newrecon_amd.ostr import load_raw, save_raw
    a = np.linspace(-0.1, 0.1, 9).astype(np.float32)
    a.tofile(os.path.join(tmp_path, "ang"))
    np.testing.assert_array_equal(load_angles(os.path.join(tmp_path, "ang"), 9), a)
    with pytest.raises(DatafileError):
        load_angles(os.path.join(tmp_path, "ang"), 10)
    g = Geometry.half_cone(4, 3, 2, 1, 1, 1, 8, 6, 1, 1)
    v = torch.arange(24, dtype=torch.float32).reshape(2, 3, 4)
    save_raw(os.path.join(tmp_path, "v.raw"), v)
    np.testing.assert_array_equal(load_raw(os.path.join(tmp_path, "v.raw"), g), v.numpy())
    with pytest.raises(ValueError):
        load_raw(os.path.join(tmp_path, "v.raw"), Geometry.half_cone(4, 3, 3, 1, 1, 1, 8, 6, 1, 1))


@pytest.mark.gpu
def test_cli_run_matches_oracle(tmp_path, small_case, oracle_lib):
    """Tomo_OSTR <datafile> equivalent: Lange prior, gamma = ray length,
    scatter r > 0, outputs every 2 iterations."""
    from newrecon_amd.ostr_cli import main
    s, x, y, b = small_case
    rng = np.random.default_rng(21)
    r = (rng.random(y.shape) * 5).astype(np.float32)
    path = write_run(str(tmp_path), y, b, r, s.angles)
    assert main([path]) == 0
    out2 = np.fromfile(os.path.join(tmp_path, "out", "recon_1"), np.float32)
    out4 = np.fromfile(os.path.join(tmp_path, "out", "recon_3"), np.float32)
    ref, snaps = oracle_lib.run(s.geom, s.angles, y, b, r=r, n_subsets=3, n_iters=4, beta=1e3,
                                delta=5e-4, potential=1, gamma_mode=0, guard=False,
                                snapshots=True)
    for got, want in ((out2, snaps[1]), (out4, snaps[3])):
        rel = np.linalg.norm(got - want.ravel()) / np.linalg.norm(want)
        assert rel <= 1e-4, rel
