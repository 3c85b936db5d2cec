"""The reference's output and restart files (sqdft_amd.files) and the
drivers' use of them: layouts checked line by line against literal text
written from the reference's printf formats (src/readfiles.cpp:1230-1646,
src/md.cpp:22-50), the rank-block order of .restart against a loop over
ranks written as PrintSCF / Processor_domain walk them, round trips, and an
MD run continued from its .restartMD file."""
import os

import numpy as np
import pytest

import helpers as H
from sqdft_amd import files, md

from test_md_host import _SCF


def rank_order_by_loops(rho, n, nproc):
    # Processor_domain (src/initialize.cpp:232-258): rank count runs over
    # x-blocks, then y, then z (innermost); PrintSCF (src/readfiles.cpp:1336-1352)
    # stores each rank's block k, j, i with i fastest
    p = round(nproc ** (1 / 3))
    b = n // p
    g = np.asarray(rho).reshape(n, n, n)          # [z][y][x]
    out = []
    for cx in range(p):
        for cy in range(p):
            for cz in range(p):
                for k in range(b):
                    for j in range(b):
                        for i in range(b):
                            out.append(g[cz * b + k, cy * b + j, cx * b + i])
    return np.array(out)


@pytest.mark.parametrize("n,nproc", [(4, 1), (4, 8), (6, 8), (6, 27), (8, 64)])
def test_restart_rank_order(n, nproc, tmp_path):
    rho = np.random.Generator(np.random.PCG64(n * 100 + nproc)).random(n ** 3)
    want = rank_order_by_loops(rho, n, nproc)
    assert np.array_equal(files.density_to_rank_order(rho, n, nproc), want)
    assert np.array_equal(files.density_from_rank_order(want, n, nproc), rho)
    p = str(tmp_path / "x.restart")
    files.write_restart(p, rho, n, nproc)
    lines = open(p).read().split("\n")
    assert lines[0] == str(nproc)
    assert lines[1] == "%.14f" % want[0] and lines[n ** 3] == "%.14f" % want[-1]
    back = files.read_restart(p, n)
    assert np.max(np.abs(back - rho)) <= 5e-15
    assert np.array_equal(files.read_restart(p, n, nproc), back)


def test_restart_errors(tmp_path):
    rho = np.zeros(4 ** 3)
    with pytest.raises(ValueError):
        files.write_restart(str(tmp_path / "a"), rho, 4, 4)        # not a cube of ranks
    with pytest.raises(ValueError):
        files.write_restart(str(tmp_path / "a"), np.zeros(5 ** 3), 5, 8)   # 5 nodes over 2 ranks
    p = str(tmp_path / "b.restart")
    files.write_restart(p, rho, 4, 8)
    with pytest.raises(ValueError):
        files.read_restart(p, 4, nproc=1)     # written by another rank count
    with pytest.raises(ValueError):
        files.read_restart(p, 5)              # wrong grid


def test_atoms_file_matches_the_library_reader():
    for d, stem, atoms in (("Al", "test_Al", "Al_eq.atoms"), ("LiH", "test_LiH", "LiH.atoms")):
        types, frac = files.read_atoms_file(os.path.join(H.GOLDEN, d, atoms))
        S = (H.test_al if d == "Al" else H.test_lih)()[0]
        xyz, typ = S.atoms()
        assert [t.natm for t in types] == [int(np.sum(typ == i)) for i in range(len(types))]
        assert [t.Z for t in types] == [S.type_info(i)["Z"] for i in range(len(types))]
        assert [t.lloc for t in types] == [S.type_info(i)["lloc"] for i in range(len(types))]
        # the library maps positions into the cell; the file's are inside already
        assert np.allclose(np.concatenate([t.xyz for t in types]), xyz, atol=1e-12)
        S.close()
    types, _ = files.read_atoms_file(os.path.join(H.GOLDEN, "LiH", "LiH.atoms"))
    assert (types[0].symbol, types[0].mass, types[0].rb) == ("Li", 6.941, 8.0)
    assert (types[1].symbol, types[1].mass, types[1].natm) == ("H", 1.00794, 4)


def test_aout_layout_and_round_trip(tmp_path):
    t = files.TypeMeta("Al", 3, 26.9815385, 2, 1, 6.0, "../../pseudopotentials/psd_Al.pot")
    xyz = np.array([[0.0, 0.5, 1.0], [3.74, 3.49, 0.37]])
    f = np.array([[1e-3, -2.5e-2, 0.0], [-1e-3, 2.5e-2, 0.0]])
    p = str(tmp_path / "Al.aout")
    files.write_aout(p, [t], xyz, f)
    want = ("2\nn_typ 1\nfrac_coord 0\n\nAtoms\n"
            "Al 3 26.981538 2 1 6.000000\n../../pseudopotentials/psd_Al.pot\n"
            "0.0000000000000000 0.5000000000000000 1.0000000000000000\n"
            "3.7400000000000002 3.4900000000000002 0.3700000000000000\n"
            " \nAtomic Forces (Ha/Bohr): \n \nAl\n"
            "0.00100000000000 -0.02500000000000 0.00000000000000\n"
            "-0.00100000000000 0.02500000000000 0.00000000000000\n")
    assert open(p).read() == want
    r = files.read_aout(p)
    assert np.array_equal(r["xyz"], np.array([[0.0, 0.5, 1.0], [3.7400000000000002, 3.4900000000000002, 0.37]]))
    assert np.array_equal(r["forces"], f)
    assert (r["types"][0].symbol, r["types"][0].natm, r["types"][0].psd_path) == ("Al", 2, t.psd_path)
    # fractional coordinates: positions divided by the lattice constant, with the note line
    files.write_aout(p, [t], xyz, f, frac_coord=1, latconst=7.78)
    lines = open(p).read().split("\n")
    assert lines[3] == "Note: Atom positions below are in fractional coordinates. "
    assert lines[9] == "%.16f %.16f %.16f" % tuple(xyz[1] / 7.78)


def test_restart_md_layout_and_round_trip(tmp_path):
    st = files.MDState(3, np.array([[1e-4, -2e-4, 0.0]]), np.array([[0.01, 0.0, -0.02]]),
                       np.array([[1.5, 2.0, 7.25]]), np.arange(8) * 0.5)
    p = str(tmp_path / "x.restartMD")
    files.write_restart_md(p, st)
    want = ("MDstepCount\n3\nMD velocities\n0.000100000000000 -0.000200000000000 0.000000000000000\n"
            "MD forces\n0.010000000000000 0.000000000000000 -0.020000000000000\n"
            "Atom positions\n1.500000000000000 2.000000000000000 7.250000000000000\n"
            "MD stats\n0.000000000000000 0.500000000000000\n1.000000000000000 1.500000000000000\n"
            "2.000000000000000 2.500000000000000\n3.000000000000000 3.500000000000000\n")
    assert open(p).read() == want
    r = files.read_restart_md(p, 1)
    assert r.step_count == 3
    for a, b in ((r.vels, st.vels), (r.forces, st.forces), (r.xyz, st.xyz), (r.sums, st.sums)):
        assert np.array_equal(a, b)
    with pytest.raises(ValueError):      # no MD stats section
        files.read_restart_md(_truncated(p, tmp_path), 1)


def _truncated(p, tmp_path):
    q = str(tmp_path / "cut.restartMD")
    with open(p) as f, open(q, "w") as g:
        g.write("".join(f.readlines()[:4]))
    return q


def test_vout_and_trajectory(tmp_path):
    v = np.array([[3.0, 4.0, 0.0], [0.0, 0.0, 1e-3]])
    p = str(tmp_path / "x.vout")
    files.write_vout(p, 1, v)
    files.write_vout(p, 2, 2 * v)
    assert open(p).read() == "1\n5.000000\n0.001000\n2\n10.000000\n0.002000\n"
    q = str(tmp_path / "Position.txt")
    files.append_trajectory(q, 1, 10, [7.78, 7.78, 7.78], [[0.0, 3.89, 1e-5]])
    files.append_trajectory(q, 2, 10, [7.78, 7.78, 7.78], [[0.5, 3.8900001, 12345678.0]])
    assert open(q).read() == ("Trajectory Data\n---------------.\n\n10\n7.78 7.78 7.78\n1\n\n"
                              " 0  3.89  1e-05\n\n 0.5  3.89  1.23457e+07\n\n")


def test_nve_continues_from_its_restart_file(tmp_path):
    # a harmonic stand-in SCF: six steps in one run equal three steps, a
    # .restartMD file, and three more steps from it
    x0 = [[5.0, 5.0, 5.0], [9.0, 9.0, 9.0]]
    full = md.NVE(_SCF(x0), masses=[1.0], time_step=0.5, T_ion=3000.0)
    ref = full.run(6)
    p = str(tmp_path / "run.restartMD")
    tr = str(tmp_path / "Position.txt")
    a = md.NVE(_SCF(x0), masses=[1.0], time_step=0.5, T_ion=3000.0, prnt_md=p, trajectory=tr, max_steps=6)
    a.run(3)
    st = files.read_restart_md(p, 2)
    assert st.step_count == 3
    b = md.NVE(_SCF(x0), masses=[1.0], time_step=0.5, T_ion=3000.0, restart_md=p, prnt_md=p, max_steps=6)
    recs = b.run(3)
    assert [r["md_step"] for r in recs] == [4, 5, 6]
    for r, q in zip(recs, ref[3:]):
        assert np.allclose(r["positions"], q["positions"], rtol=0, atol=1e-12)
        assert r["TE"] == pytest.approx(q["TE"], rel=1e-12)
    mb, mf = b.means(), full.means()
    for k in ("TE", "PE", "KE", "T"):
        assert mb[k][0] == pytest.approx(mf[k][0], rel=1e-12)
    assert files.read_restart_md(p, 2).step_count == 6
    with pytest.raises(ValueError):     # every step already done
        md.NVE(_SCF(x0), masses=[1.0], restart_md=p, max_steps=6)
    assert open(tr).read().count("Trajectory Data") == 1
