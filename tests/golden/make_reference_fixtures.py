"""Copy the data fixtures the reference ships in run/ into tests/golden/reference_run/.

Sources (data files of the reference repository, not code):
  run/grid.dat    Fortran unformatted: nx,ny,nzp; x0,xl,xlen,ylen,zlen; xpts; ypts; zpts; gf
                  (written by write_grid, io.f90:110-130) for the run/input.dat case
  run/paral.dat   Fortran unformatted: nzp; ubar, ubarp, ubarpp, wbar, wbarp, wbarpp, tbar, tbarp, tbarpp
                  (write_paral, io.f90:251-275)
  run/output.txt  log of the same run: the 31 eigenvalues of d2(2:nz,2:nz) printed by
                  diagonalize_matrix (helmholtz.f90:112-120), the Blasius iteration count,
                  the initial dt and the maxcalc lines (navierstokes.f90:494-506)
  run/q00019.dat  mpi_snapshot file (mpi_io.f90:339-497) of a 224 x 24 x 33 run, stored whole
                  as float32 in q00019.npz
  run/prev_q00001.dat  the snapshot of step 10 of the run/input.dat case (Blasius base flow,
                  perturbations ~1e-17): its header and the y = 0 planes of u and temp
  run/energy.dat, run/output_lintest.txt  the linear-stability sweep (write_linstab_results,
                  orrsomm.f90:63-107): per (Re_d, alpha*) omega_i, <dE/dt>, % error, and the last
                  omega of the Orr-Sommerfeld iteration printed to 10 digits (orrsomm.f90:428)
Run from the repository root:  python tests/golden/make_reference_fixtures.py
"""

import json
import os
import shutil

import numpy as np

REF = "/root/reference/run"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_run")


def main():
    os.makedirs(OUT, exist_ok=True)
    for f in ("grid.dat", "paral.dat"):
        shutil.copyfile(os.path.join(REF, f), os.path.join(OUT, f))
    txt = open(os.path.join(REF, "output.txt")).read()
    sect = txt.split("eigenvalues =")[1].split("max iter")[0]
    ev = [float(line.split(",")[0]) for line in sect.strip().splitlines() if line.strip()]
    iters = int(txt.split("blasfl: converged at iter=")[1].split()[0])
    dt = float(txt.split("based input cfl, grid, and initial dt =")[1].split()[0])
    div = [float(x.split(",")[0]) for x in txt.split("divergence avg error =")[1:]]
    maxu = [float(x.split()[0]) for x in txt.split("Max|u+ubar|=")[1:]]
    cflx = [float(x.split()[0]) for x in txt.split("Max. x-CFL#:")[1:]]
    rho = [float(x.split()[0]) for x in txt.split("Max|rho|=")[1:]]
    meta = {
        "case": "run/input.dat: nx=224 ny=24 nz=32 mapping=2 xlen=333.3333 ylen=20 zlen=33.3333 x0=8.3333 "
                "u0=1 xnu=0.001 blfac=8 cfl=0.1 bctype=0 sgsflag=0 dealiasflag=2",
        "eigenvalues_print_precision": "E14.7",
        "eigenvalues": ev,
        "blasfl_iterations": iters,
        "initial_dt": dt,
        "divergence_avg_error": div,
        "max_u_plus_ubar": maxu,
        "max_x_cfl": cflx,
        "max_rho": rho,
    }
    with open(os.path.join(OUT, "output_txt.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    q19 = np.fromfile(os.path.join(REF, "q00019.dat"), dtype="<f4")
    np.savez_compressed(os.path.join(OUT, "q00019.npz"), record=q19)
    q1 = np.fromfile(os.path.join(REF, "prev_q00001.dat"), dtype="<f4")
    nx, ny, nzp = (int(v) for v in q1[3:6])
    body = q1[7:].reshape(5, ny, nzp, nx)
    np.savez_compressed(os.path.join(OUT, "prev_q00001_y0.npz"), header=q1[:7], u_y0=body[0, 0], temp_y0=body[3, 0])
    rows = []
    for line in open(os.path.join(REF, "energy.dat")).read().splitlines()[1:]:
        v = [float(x) for x in line.replace(",", " ").split()]
        if len(v) >= 5:
            rows.append(dict(re_d=v[0], alphastar=v[1], omega_i=v[2], dedt=v[3], err_pct=v[4]))
    cases = []
    cur = None
    for line in open(os.path.join(REF, "output_lintest.txt")):
        if line.startswith(" Re_d ="):
            parts = line.replace("(", " ").replace(")", " ").replace(",", " ").split()
            cur = dict(re_d=float(parts[2]), x0=float(parts[5]), alphastar=float(parts[8]),
                       cguess=float(parts[11]))
            cases.append(cur)
        elif cur is not None and line.strip().startswith("100,"):
            v = [float(x) for x in line.replace(",", " ").split()]
            cur["omega"] = [v[1], v[2]]
    with open(os.path.join(OUT, "linstab.json"), "w") as fh:
        json.dump({"energy_dat": rows, "orrsomm_omega": [c for c in cases if "omega" in c]}, fh, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
