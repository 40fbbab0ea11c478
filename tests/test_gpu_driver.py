"""The reference-style driver lsg_solve (GMRES/GMRES.cpp:main restated): its
12 positional inputs, the result/ files (6-digit Phi, grid points, total field,
leaf boundaries), and the field reconstruction u = u_inc + V Phi checked
against the Lippmann-Schwinger consistency Phi = -k^2 q u."""
import os
import re
import subprocess

import numpy as np
import pytest

import lippmann_schwinger_solver_amd as L
from tests import helpers as H

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(L.LIB_PATH), "lsg_solve")


def read_phi(path):
    vals = []
    for line in open(path):
        m = re.match(r"\(([^,]+),([^)]+)\)", line.strip())
        vals.append(complex(float(m.group(1)), float(m.group(2))))
    return np.array(vals)


# GMRES/Readme.md:64-101: "Err in GMRES" after each sweep entry of the README run
README_ARGS = ["16", "6", "5", "0.5", "40.0", "1", "6", "8", "10", "1", "5", "0"]
README_ERR = [4.59412e-06, 1.40294e-15, 1.30732e-15, 1.39836e-15, 1.40164e-15,
              1.28443e-10, 8.96364e-16, 7.94185e-16, 8.37373e-16, 7.82761e-16]
README_TIME = [0.277782, 1.47135, 2.90013, 6.22192, 19.1974, 0.751761, 2.63477, 5.29393, 11.7579, 38.9819]


def test_driver_reproduces_readme_run(tmp_path):
    """lsg_solve with the README's own inputs (adaptive tree, kappa 40, TOL_POW 8):
    the printed tree and the residual of every GMRES sweep entry. The two
    entries above machine precision agree to all six printed digits; the rest
    are at the FP64 floor like the README's (1e-15)."""
    print("lsg_solve " + " ".join(README_ARGS), flush=True)
    r = subprocess.run([EXE] + README_ARGS, cwd=tmp_path, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "nLevels: 5" in out and "level_LFR: 4" in out and "Number of particles: 5328" in out
    errs = [float(v) for v in re.findall(r"Err in GMRES: (\S+)", out)]
    times = [float(v) for v in re.findall(r"Time taken by GMRES Solver \d+,\d+: (\S+)", out)]
    assert len(errs) == 10 and len(times) == 10
    for e, ref, t, tref in zip(errs, README_ERR, times, README_TIME):
        print(f"residual {e:.5e} (README {ref:.5e}), {t:.4f} s (README {tref} s)", flush=True)
    assert float(f"{errs[0]:.5e}") == README_ERR[0]
    assert float(f"{errs[5]:.5e}") == README_ERR[5]
    assert all(e < 1e-13 for i, e in enumerate(errs) if i not in (0, 5))
    d = tmp_path / "result" / "result_0_6_5_40_10"
    for n in ("Phi_1_10", "Phi_1_500", "Phi_2_10", "Phi_2_50", "gridPointsX", "solutionR", "leftRightBoundary"):
        assert (d / n).exists(), n
    lr = np.loadtxt(d / "leftRightBoundary")
    assert lr.shape == (148, 2)


def test_driver_result_files(tmp_path):
    # uniform tree of depth 3 (optional argv[16] = 0), kappa 10, Gaussian
    # contrast (Qchoice 0), TOL_POW 10, m = 50, restart 1
    args = ["16", "6", "5", "0.5", "10.0", "1", "6", "10", "50", "1", "3", "0", "0", "0", "5", "0"]
    print("lsg_solve " + " ".join(args), flush=True)
    r = subprocess.run([EXE] + args, cwd=tmp_path, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    assert "Err in GMRES" in r.stdout
    d = tmp_path / "result" / "result_0_6_5_10_50"
    names = sorted(os.listdir(d))
    for n in ("Phi_1_10", "Phi_1_50", "Phi_2_50", "gridPointsX", "gridPointsY", "solutionR", "solutionI",
              "leftRightBoundary", "bottomTopBoundary"):
        assert n in names, n
    S = L.DFMM(H.inputs(depth=3, kappa=10.0))
    S.FMMTreeInitialisation()
    N = S.N
    phi6 = read_phi(d / "Phi_1_50")
    phi = np.fromfile(d / "Phi_1_50.f64", dtype=np.complex128)
    assert phi.size == N and phi6.size == N
    # the text file holds the sidecar at 6 significant digits
    assert np.max(np.abs(phi6 - phi)) <= 1e-5 * np.max(np.abs(phi))
    b = S.rhs()
    assert H.relerr(S.MatVecProduct(phi), b) < 1e-8          # GMRES(50), one restart, at depth 3
    gx = np.loadtxt(d / "gridPointsX")
    gy = np.loadtxt(d / "gridPointsY")
    assert np.allclose(gx, S.gridPoints[:, 0], atol=1e-5) and np.allclose(gy, S.gridPoints[:, 1], atol=1e-5)
    # total field written from the last sweep entry (the (2,500) run, file Phi_2_50)
    u = np.loadtxt(d / "solutionR") + 1j * np.loadtxt(d / "solutionI")
    phi_last = np.fromfile(d / "Phi_2_50.f64", dtype=np.complex128)
    V = L.DFMM(H.inputs(depth=3, kappa=10.0, find_phi=0))
    V.FMMTreeInitialisation()
    u_ref = S.incident() + V.MatVecProduct(phi_last)
    assert H.relerr(u, u_ref) < 1e-5                         # 6-digit text
    # Lippmann-Schwinger consistency: Phi = -k^2 q u (hybridTrueValidation_gmres.cpp:227-232)
    qk2 = S.contrast_k2()
    cons = H.relerr(-qk2 * u_ref, phi_last)
    print(f"consistency |Phi + k^2 q u| / |Phi| = {cons:.2e}", flush=True)
    assert cons < 1e-4      # V alone and I + DV carry different NCA errors (about 6e-6 at depth 3)
    lr = np.loadtxt(d / "leftRightBoundary")
    assert lr.shape == (64, 2) and np.all(lr[:, 1] > lr[:, 0])
