"""The run drivers (qed2d.run, counterpart of main/hmc-nf2.jl, main/rhmc.jl,
main/MC.jl): input schema of extra/*.in and the statistics / config file
conventions on the CPU; whole runs, and a continued run reproducing an
uninterrupted one, on the GPU."""
import json
import os

import pytest

NF2_IN = {"Run": {"user": "u", "host": "h", "name": "t", "seed": 77},
          "Lattice": {"size": [8, 6], "beta": 4.0},
          "HMC": {"tau": 1.0, "ns": 10, "nthm": 0, "ntraj": 4, "Qzero": False, "mass": 0.3}}
RHMC_IN = {"Run": {"user": "u", "host": "h", "name": "t", "seed": 5},
           "Lattice": {"size": [8, 8], "beta": 5.0},
           "HMC": {"tau": 1.0, "ns": 10, "nthm": 0, "ntraj": 2, "Qzero": False},
           "RHMC": {"masses": [0.3, 0.3], "n_rhmc": [4, 4], "r_a_g5Dw_sqr": [0.05, 0.05],
                    "r_b_g5Dw_sqr": [40.0, 40.0]}}
MC_IN = {"Run": {"user": "u", "host": "h", "name": "t", "seed": 1234, "measurements": 4},
         "Lattice": {"size": [16, 16], "beta": 4.0},
         "HMC": {"ns": 10, "eps": 0.1, "ntraj": 3, "nthm": 2, "Qzero": False}}


def _write(tmp_path, name, d):
    p = tmp_path / name
    p.write_text(json.dumps(d))
    return str(p)


def test_read_options(tmp_path):
    from qed2d import run as R
    o = R.read_options(_write(tmp_path, "nf2.in", NF2_IN), "hmc-nf2")
    assert o["size"] == (8, 6) and o["mass"] == 0.3 and o["eps"] == 0.1 and o["seed"] == 77
    o = R.read_options(_write(tmp_path, "rhmc.in", RHMC_IN), "rhmc")
    assert o["masses"] == [0.3, 0.3] and o["n_rhmc"] == [4, 4]
    assert abs(o["r_a"][0] - 0.05 ** 0.5) < 1e-15 and abs(o["r_b"][1] - 40.0 ** 0.5) < 1e-15
    o = R.read_options(_write(tmp_path, "mc.in", MC_IN), "mc")
    assert o["eps"] == 0.1 and o["nthm"] == 2
    d = json.loads(json.dumps(NF2_IN))
    d["Run"]["seed"] = 0
    assert R.read_options(_write(tmp_path, "z.in", d), "hmc-nf2")["seed"] > 0


def test_config_name():
    import qed2d as Q
    prm = Q.LattParm((128, 64), 11.25)
    assert Q.config_name(prm, 0.2, 300) == "config_128_64_b11.25_m0.2_n300"


@pytest.mark.gpu
def test_nf2_run_and_continuation(tmp_path):
    from qed2d import run as R
    inp = _write(tmp_path, "nf2.in", NF2_IN)
    a = str(tmp_path / "a")
    rows = R.run("hmc-nf2", inp, a, 0, save_every=2, tol=1e-20)
    assert len(rows) == 4
    lines = open(os.path.join(a, "statistics.txt")).read().splitlines()
    assert lines[0] == "step,accepted,plaquette,top_charge,acc_rate"
    assert [ln.split(",")[0] for ln in lines[1:]] == ["1", "2", "3", "4"]
    assert os.path.exists(os.path.join(a, "configs", "config_8_6_b4.0_m0.3_n2"))
    # continue from configuration 2: trajectories 3 and 4 again
    b = str(tmp_path / "b")
    os.makedirs(os.path.join(b, "configs"))
    for f in ("config_8_6_b4.0_m0.3_n2", "config_8_6_b4.0_m0.3_n2.rng.json"):
        with open(os.path.join(a, "configs", f)) as src, open(os.path.join(b, "configs", f),
                                                             "w") as dst:
            dst.write(src.read())
    open(os.path.join(b, "statistics.txt"), "w").write("".join(ln + "\n" for ln in lines[:3]))
    d = json.loads(json.dumps(NF2_IN))
    d["HMC"]["ntraj"] = 2
    rows_b = R.run("hmc-nf2", _write(tmp_path, "nf2b.in", d), b, 2, save_every=2, tol=1e-20)
    # the acceptance rate restarts with the continued run (as the reference's
    # acc vector does); step, accept, plaquette and charge must agree
    assert [r.split(",")[:4] for r in rows_b] == [r.split(",")[:4] for r in rows[2:]]


@pytest.mark.gpu
def test_rhmc_and_mc_runs(tmp_path):
    from qed2d import run as R
    rows = R.run("rhmc", _write(tmp_path, "r.in", RHMC_IN), str(tmp_path / "r"), 0, tol=1e-20)
    assert len(rows) == 2 and all(len(r.split(",")) == 8 for r in rows)
    head = open(str(tmp_path / "r" / "statistics.txt")).readline()
    assert head == R.RHMC_HEADER
    w = float(rows[0].split(",")[4])
    assert 0.5 < w < 2.0
    rows = R.run("mc", _write(tmp_path, "m.in", MC_IN), str(tmp_path / "m"), 0)
    assert len(rows) == 3


MF_IN = {"Run": {"user": "u", "host": "h", "name": "t", "seed": 21, "measurements": 2},
         "Lattice": {"size": [64, 64], "beta": 11.25},
         "HMC": {"ns": 4, "eps": 0.05, "ntraj": 1, "Qzero": False}}


def test_err_avg_blocks():
    """err_avg (main/master_field.jl:43-57): 16 x 16 block averages, their mean
    and sample standard deviation / 16; a field constant on each block with
    block (b1, b2) = b1 + 16 b2 has mean 127.5 and the spread of 0..255."""
    import numpy as np
    from qed2d import run as R
    L1, L2 = 64, 48
    d = np.zeros((L2, L1))
    for i2 in range(L2):
        for i1 in range(L1):
            d[i2, i1] = (i1 // (L1 // 16)) + 16 * (i2 // (L2 // 16))
    m, e = R.err_avg(d)
    assert m == 127.5
    assert abs(e - np.std(np.arange(256.0), ddof=1) / 16) < 1e-12
    m, e = R.err_avg(np.full((32, 32), 0.75))
    assert m == 0.75 and e == 0.0


def test_master_field_options(tmp_path):
    from qed2d import run as R
    o = R.read_options(_write(tmp_path, "mf.in", MF_IN), "master-field")
    assert o["measurements"] == 2 and o["eps"] == 0.05 and o["ns"] == 4 and o["size"] == (64, 64)


@pytest.mark.gpu
def test_master_field_run(tmp_path, O):
    """Growth 16 -> 32 -> 64 by unfold_fld, two measurement rounds and a last
    save: file sizes and layout, and the last saved densities against the
    oracle's plaquette and charge densities of the final links."""
    import numpy as np
    from qed2d import run as R
    out = str(tmp_path / "mf")
    U, prm, rows = R.master_field(_write(tmp_path, "mf.in", MF_IN), out)
    assert prm.iL == (64, 64) and tuple(U.shape) == (2, 64, 64)
    meta = json.load(open(os.path.join(out, "master_field.json")))
    assert meta["L1"] == 64 and meta["measurements"] == 2
    raw = np.fromfile(os.path.join(out, "measurements.f64"), dtype="<f8")
    assert raw.size == 3 * 2 * 64 * 64
    last = raw[-2 * 64 * 64:].reshape(2, 64, 64)  # [density][i1][i2]
    Uh = U.cpu().numpy()
    assert np.max(np.abs(last[0].T - O.plaquette_field(Uh))) < 1e-13
    assert np.max(np.abs(last[1].T - O.qtop_field(Uh))) < 1e-13
    lines = open(os.path.join(out, "master_field.txt")).read().splitlines()
    assert lines[0] == "save,plaquette,plaquette_err,top_charge" and len(lines) == 4
    m, _ = R.err_avg(last[0].T)
    assert abs(float(lines[-1].split(",")[1]) - m) < 1e-15


def _rows(path):
    return [ln.split(",") for ln in open(path).read().splitlines()[1:]]


@pytest.mark.gpu
def test_driver_rows_match_oracle(tmp_path, O):
    """Each statistics row (main/hmc-nf2.jl:129-140, main/rhmc.jl:141-155,
    main/MC.jl:120-135) against the oracle run on the same inputs and
    streams: trajectory k of a run draws from (seed, chain 0, trajectory k);
    the rhmc row's reweighting factor takes one trajectory index per fermion
    and the power method's start vector the next (flavour 15).  Accept/reject
    equal, plaquette and Q to 1e-10, the reweighting factor and lambda_max to
    1e-8 relative (lambda_min: see below)."""
    import numpy as np
    from dense import q_dense
    from qed2d import run as R
    # hmc-nf2: cold start, leapfrog, CG tol 1e-20
    rows = R.run("hmc-nf2", _write(tmp_path, "nf2.in", NF2_IN), str(tmp_path / "a"), 0,
                 tol=1e-20)
    o = R.read_options(_write(tmp_path, "nf2b.in", NF2_IN), "hmc-nf2")
    L1, L2 = o["size"]
    U = np.ones((2, L2, L1), complex)
    p = O.latt(L1, L2, o["beta"])
    nacc = 0
    for t, row in enumerate(_rows(str(tmp_path / "a" / "statistics.txt"))):
        _, _, a, _, rc = O.hmc_nf2(U, o["mass"], o["eps"], o["ns"], 10000, 1e-20, p, o["seed"], t)
        assert rc == 0
        nacc += a
        assert int(row[0]) == t + 1 and int(row[1]) == a, (t, row)
        assert abs(float(row[2]) - O.plaquette(U)) < 1e-10
        assert abs(float(row[3]) - O.qtop(U)) < 1e-10
        assert abs(float(row[4]) - nacc / (t + 1)) < 1e-15
    assert len(rows) == o["ntraj"]
    # mc: pure gauge OMF4 after nthm thermalisation trajectories
    R.run("mc", _write(tmp_path, "m.in", MC_IN), str(tmp_path / "m"), 0)
    o = R.read_options(_write(tmp_path, "mb.in", MC_IN), "mc")
    L1, L2 = o["size"]
    U = np.ones((2, L2, L1), complex)
    p = O.latt(L1, L2, o["beta"])
    for t in range(o["nthm"]):
        O.hmc_pure(U, o["eps"], o["ns"], p, o["seed"], t)
    for j, row in enumerate(_rows(str(tmp_path / "m" / "statistics.txt"))):
        a = O.hmc_pure(U, o["eps"], o["ns"], p, o["seed"], o["nthm"] + j)[2]
        assert int(row[1]) == a, (j, row)
        assert abs(float(row[2]) - O.plaquette(U)) < 1e-10
        assert abs(float(row[3]) - O.qtop(U)) < 1e-10
    # rhmc: two fermions, leapfrog, multi-shift solves on the GPU against the
    # oracle's sequential ones
    R.run("rhmc", _write(tmp_path, "r.in", RHMC_IN), str(tmp_path / "r"), 0, tol=1e-20)
    o = R.read_options(_write(tmp_path, "rb.in", RHMC_IN), "rhmc")
    L1, L2 = o["size"]
    U = np.ones((2, L2, L1), complex)
    p = O.latt(L1, L2, o["beta"])
    zs = [O.zolotarev(n, a, b) for n, a, b in zip(o["n_rhmc"], o["r_a"], o["r_b"])]
    import qed2d as Q
    rps = Q.get_rhmc_params(o["n_rhmc"], o["r_a"], o["r_b"])  # (reweighting N, Taylor order)
    t = 0
    for row in _rows(str(tmp_path / "r" / "statistics.txt")):
        _, _, a, rc = O.hmc_rhmc(U, o["masses"], zs, o["eps"], o["ns"], 10000, 1e-20, p,
                                 o["seed"], t)[:4]
        t += 1
        assert rc == 0 and int(row[1]) == a, row
        assert abs(float(row[2]) - O.plaquette(U)) < 1e-10
        assert abs(float(row[3]) - O.qtop(U)) < 1e-10
        w = 1.0
        for m, z, rp in zip(o["masses"], zs, rps):
            w *= O.reweighting_factor(U, m, z, rp.reweighting_N, rp.reweighting_Taylor, 10000,
                                      1e-20, o["seed"], 0, t)[0]
            t += 1
        assert abs(float(row[4]) - w) < 1e-8 * abs(w)
        b0 = O.pseudofermion(L1, L2, o["seed"], 0, t, 15)
        t += 1
        lmin, lmax = O.power_method(U, b0, o["masses"][0], 5000)
        assert abs(float(row[6]) - lmax) < 1e-8 * abs(lmax)
        # lambda_min: its loop grows the low modes from the rounding noise the
        # lambda_max loop leaves (device-specific), and 5000 steps leave it a
        # Rayleigh quotient above the lowest eigenvalue: GPU and oracle to 1 %,
        # neither below the dense spectrum's minimum
        ev = np.linalg.eigvalsh(q_dense(U, o["masses"][0]) @ q_dense(U, o["masses"][0]))
        assert abs(float(row[5]) - lmin) < 1e-2 * abs(lmin)
        assert float(row[5]) > ev[0] * (1 - 1e-9) and lmin > ev[0] * (1 - 1e-9)
