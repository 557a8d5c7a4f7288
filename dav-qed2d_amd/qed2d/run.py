"""Run drivers: the counterpart of the reference's main/hmc-nf2.jl,
main/rhmc.jl and main/MC.jl, with their JSON input files (extra/*.in) and
their outputs, on this package.

    python -m qed2d.run hmc-nf2 -i extra/hmc-nf2.in --savedir runs/a/ --cont 0
    python -m qed2d.run rhmc    -i extra/rhmc.in    --savedir runs/b/ --cont 0
    python -m qed2d.run mc      -i extra/MC.in      --savedir runs/c/
    python -m qed2d.run master-field -i extra/master_field.in --savedir runs/d/

Outputs in ``savedir`` (main/hmc-nf2.jl:100-147, main/rhmc.jl:107-160):

* ``statistics.txt``: CSV, header ``step,accepted,plaquette,top_charge,acc_rate``
  (hmc-nf2, mc) or ``step,accepted,plaquette,top_charge,reweight_factor,
  lambda_min,lambda_max,acc_rate`` (rhmc), one line per trajectory;
* ``log.txt``: the run's parameters;
* ``configs/config_{L1}_{L2}_b{beta}_m{mass}_n{i}``: the links every 100
  (hmc-nf2) / 10 (rhmc) trajectories, in the reference's text format
  (src/QED2dLoadSave.jl);
* ``configs/<config name>.rng.json``: the random-stream position (seed,
  chain, next trajectory) at each saved configuration, which ``--cont``
  resumes exactly.  The reference reseeds on continuation
  (main/hmc-nf2.jl:40-46), so its continued runs do not reproduce an
  uninterrupted one; this driver's do.

``Run.seed`` 0 draws a random seed, as the reference.  The reference's MC.jl
writes BDIO files; ``mc`` writes ``statistics.txt`` instead (BDIO is not part of
this package).

``master-field`` (main/master_field.jl): pure-gauge HMC grown from a cold 16 x 16
lattice by ``unfold_fld`` doublings (``ntraj`` trajectories at each size) up to
the target size, then ``measurements`` rounds of (save the per-site densities,
``ntraj`` trajectories) and a last save.  The reference's BDIO records become
``measurements.f64``: per save, the plaquette density then the topological-charge
density, each as L1 rows of L2 little-endian doubles (the reference writes
``dfoo[i,:]`` for i = 1..L1); ``master_field.json`` describes the file and
``master_field.txt`` holds, per save, the plaquette mean and error from 16 x 16
blocks (``err_avg``, main/master_field.jl:43-57) and Q.  The reference seeds
with 1234 whatever the input says (main/master_field.jl:69-70); this driver uses
``Run.seed``.
"""
import argparse
import json
import os
import random
import sys
import time

import qed2d as Q

NF2_HEADER = "step,accepted,plaquette,top_charge,acc_rate\n"
RHMC_HEADER = ("step,accepted,plaquette,top_charge,reweight_factor,lambda_min,lambda_max,"
               "acc_rate\n")


def read_options(fname, mode):
    """The JSON schema of extra/hmc-nf2.in, extra/rhmc.in, extra/MC.in
    (main/hmc-nf2.jl:26-57, main/rhmc.jl:26-63, main/MC.jl:62-100)."""
    with open(fname) as f:
        s = json.load(f)
    o = {"name": s["Run"].get("name", "run"), "seed": int(s["Run"].get("seed", 0)),
         "size": (int(s["Lattice"]["size"][0]), int(s["Lattice"]["size"][1])),
         "beta": float(s["Lattice"]["beta"]), "qzero": bool(s["HMC"].get("Qzero", False)),
         "ns": int(s["HMC"]["ns"]), "ntraj": int(s["HMC"]["ntraj"]),
         "nthm": int(s["HMC"].get("nthm", 0))}
    if mode == "master-field":
        o["measurements"] = int(s["Run"].get("measurements", 0))
    if mode in ("mc", "master-field"):
        o["eps"] = float(s["HMC"]["eps"])
    else:
        o["tau"] = float(s["HMC"]["tau"])
        o["eps"] = o["tau"] / o["ns"]
    if mode == "hmc-nf2":
        o["mass"] = float(s["HMC"]["mass"])
    if mode == "rhmc":
        r = s["RHMC"]
        o["masses"] = [float(m) for m in r["masses"]]
        o["n_rhmc"] = [int(n) for n in r["n_rhmc"]]
        # the input holds the squared bounds (main/rhmc.jl:56-57)
        o["r_a"] = [float(v) ** 0.5 for v in r["r_a_g5Dw_sqr"]]
        o["r_b"] = [float(v) ** 0.5 for v in r["r_b_g5Dw_sqr"]]
    if o["seed"] == 0:
        o["seed"] = random.SystemRandom().randrange(1, 2 ** 63)
    return o


def err_avg(dens):
    """Mean and error of a density over 16 x 16 blocks (main/master_field.jl:43-57):
    the block averages' mean and their sample standard deviation / 16.  dens is
    indexed [i2, i1] (this package's layout)."""
    import numpy as np
    L2, L1 = dens.shape
    b1, b2 = L1 // 16, L2 // 16
    d = np.asarray(dens)[:16 * b2, :16 * b1]
    blks = d.reshape(16, b2, 16, b1).sum(axis=(1, 3)) / (b1 * b2)
    return float(blks.mean()), float(blks.std(ddof=1) / 16.0)


def master_field(infile, savedir, out=None, start=16):
    """The master-field run (main/master_field.jl); returns the final links,
    their lattice parameters and the rows of master_field.txt."""
    import numpy as np
    import torch
    o = read_options(infile, "master-field")
    out = out or sys.stdout
    os.makedirs(savedir, exist_ok=True)
    L1t, L2t = o["size"]
    if L1t < 128 or L2t < 128:
        print("warning: lattice is too small for a master field", file=out, flush=True)
    Q.seed(o["seed"])
    l1 = l2 = start
    prm = Q.LattParm((l1, l2), o["beta"])
    U = Q.cold_links(prm)
    acc = []
    while True:
        print(f"## lattice {l1} x {l2}", file=out, flush=True)
        for _ in range(o["ntraj"]):
            Q.HMC(U, o["eps"], o["ns"], acc, prm, None, qzero=o["qzero"])
        if 2 * l1 > L1t:
            break
        U2 = Q.unfold_fld(U, prm)
        prm.close()
        l1, l2 = 2 * l1, 2 * l2
        prm = Q.LattParm((l1, l2), o["beta"])
        U = U2
    with open(os.path.join(savedir, "master_field.json"), "w") as f:
        json.dump({"L1": l1, "L2": l2, "beta": o["beta"], "measurements": o["measurements"],
                   "records": "per save: plaquette density, then Q density; each L1 rows "
                              "of L2 float64 (little endian)"}, f)
    rows = []
    dens = torch.zeros((l2, l1), dtype=torch.float64, device=U.device)
    fdat = open(os.path.join(savedir, "measurements.f64"), "wb")
    with open(os.path.join(savedir, "master_field.txt"), "w") as ft:
        ft.write("save,plaquette,plaquette_err,top_charge\n")

        def save(j):
            Q.plaquette(dens, U, prm)
            d = dens.cpu().numpy()
            fdat.write(np.ascontiguousarray(d.T).astype("<f8").tobytes())
            pavg, perr = err_avg(d)
            Q.qtop(dens, U, prm)
            fdat.write(np.ascontiguousarray(dens.cpu().numpy().T).astype("<f8").tobytes())
            qt = Q.Qtop(U, prm)
            row = f"{j},{pavg!r},{perr!r},{qt!r}\n"
            ft.write(row)
            rows.append(row)
            print(f"measurement {j}: plaquette {pavg:.8f} +/- {perr:.2e}, Q {qt:.6f}", file=out,
                  flush=True)

        for j in range(1, o["measurements"] + 1):
            save(j)
            for _ in range(o["ntraj"]):
                Q.HMC(U, o["eps"], o["ns"], acc, prm, None, qzero=o["qzero"])
        save(o["measurements"] + 1)
    fdat.close()
    return U, prm, rows


def _mass_tag(o, mode):
    return o["mass"] if mode == "hmc-nf2" else (o["masses"][0] if mode == "rhmc" else 0.0)


def _write_state(path):
    seed, chain, traj = Q.rng_state()
    with open(path + ".rng.json", "w") as f:
        json.dump({"seed": seed, "chain": chain, "traj": traj}, f)


def _read_state(path):
    with open(path + ".rng.json") as f:
        return json.load(f)


def run(mode, infile, savedir, cont=0, save_every=None, maxiter=10000, tol=1e-16, out=None,
        precision=Q.PREC_F64):
    """Runs the driver; returns the list of statistics rows written."""
    import torch
    o = read_options(infile, mode)
    out = out or sys.stdout
    savedir = savedir if savedir.endswith(os.sep) else savedir + os.sep
    L1, L2 = o["size"]
    prm = Q.LattParm((L1, L2), o["beta"])
    mtag = _mass_tag(o, mode)
    U = Q.cold_links(prm)
    if cont:
        cfg = savedir + "configs/" + Q.config_name(prm, mtag, cont)
        Q.load_gauge(U, cfg, prm)
        st = _read_state(cfg)
        Q.seed(st["seed"], st["chain"], st["traj"])
    else:
        Q.seed(o["seed"])
    rprm = None
    if mode == "rhmc":
        rprm = Q.get_rhmc_params(o["n_rhmc"], o["r_a"], o["r_b"])
    every = save_every or (10 if mode == "rhmc" else 100)
    header = RHMC_HEADER if mode == "rhmc" else NF2_HEADER
    stat = savedir + "statistics.txt"
    log = savedir + "log.txt"
    if not cont:
        os.makedirs(savedir + "configs", exist_ok=True)
        with open(stat, "w") as f:
            f.write(header)
        with open(log, "w") as f:
            f.write(f"Start {mode} run in QED2d with beta = {o['beta']}; {L1} x {L2}; "
                    f"mass = {mtag}; ntraj = {o['ntraj']}; seed = {o['seed']}, "
                    f"nsteps={o['ns']}, eps={o['eps']}\n")
    else:
        with open(log, "a") as f:
            f.write(f"Continue {mode} run from n={cont} with beta = {o['beta']}; {L1} x {L2}; "
                    f"mass = {mtag}; ntraj = {o['ntraj']}; seed = {o['seed']}, "
                    f"nsteps={o['ns']}, eps={o['eps']}\n")
    if mode == "mc" and not cont:
        # thermalisation (main/MC.jl:112-118)
        for _ in range(o["nthm"]):
            Q.HMC(U, o["eps"], o["ns"], [], prm, None, qzero=o["qzero"])
    acc = []
    rows = []
    for i in range(1 + cont, o["ntraj"] + cont + 1):
        t0 = time.time()
        if mode == "mc":
            Q.HMC(U, o["eps"], o["ns"], acc, prm, None, qzero=o["qzero"])
        elif mode == "hmc-nf2":
            Q.HMC(U, o["mass"], o["eps"], o["ns"], acc, maxiter, tol, prm, None,
                  qzero=False, precision=precision)
        else:
            Q.HMC(U, o["masses"], o["eps"], o["ns"], acc, maxiter, tol, prm, None, rprm,
                  qzero=False)
        plaq = Q.Plaquette(U, prm)
        qt = Q.Qtop(U, prm)
        rate = sum(acc) / len(acc)
        if mode == "rhmc":
            w = Q.reweighting_factor(U, o["masses"], maxiter, tol, prm, None, rprm)
            lmin, lmax = Q.power_method(U, o["masses"][0], prm, None, iter=5000)
            row = f"{i},{acc[-1]},{plaq!r},{qt!r},{w!r},{lmin!r},{lmax!r},{rate!r}\n"
        else:
            row = f"{i},{acc[-1]},{plaq!r},{qt!r},{rate!r}\n"
        with open(stat, "a") as f:
            f.write(row)
        rows.append(row)
        if i % every == 0:
            cfg = savedir + "configs/" + Q.config_name(prm, mtag, i)
            Q.save_gauge(U, cfg, prm)
            _write_state(cfg)
        torch.cuda.synchronize()
        print(f"trajectory {i}: accepted {acc[-1]}, plaquette {plaq:.8f}, Q {qt:.6f}, "
              f"{time.time() - t0:.3f}s", file=out, flush=True)
    prm.close()
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m qed2d.run")
    ap.add_argument("mode", choices=["hmc-nf2", "rhmc", "mc", "master-field"])
    ap.add_argument("-i", required=True, help="input file")
    ap.add_argument("--savedir", default="./", help="directory for statistics, logs, configs")
    ap.add_argument("--cont", type=int, default=0,
                    help="continue from configuration number; 0 for a new run")
    ap.add_argument("--save-every", type=int, default=None)
    ap.add_argument("--mixed", action="store_true",
                    help="FP32 MD solves with FP64 reliable updates (hmc-nf2)")
    a = ap.parse_args(argv)
    if a.mode == "master-field":
        master_field(a.i, a.savedir)
        return
    run(a.mode, a.i, a.savedir, a.cont, a.save_every,
        precision=Q.PREC_MIXED if a.mixed else Q.PREC_F64)


if __name__ == "__main__":
    main()
