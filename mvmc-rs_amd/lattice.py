"""Expert-mode input generator for the benchmark workloads (a small StdFace
stand-in, crates/mvmc-stdface/src/{chain_lattice.rs,square_lattice.rs}):
periodic Hubbard chains and square lattices with t / t' hopping, on-site U,
sublattice-translation orbitals, Gutzwiller and Jastrow factors and
translation quantum-number projection.  Files follow the .def layouts the
loader reads (and the reference's parsers accept): five header lines,
index blocks, then (index, flag) optimisation pairs.

Workloads (sizes of the configurations they are named after):
  chain16    L=16 chain, t=1, U=4, 8+8 electrons, Gutzwiller (1) + nearest-
             neighbour Jastrow (1 optimised; the other pairs share a frozen
             class at 0) + 2-site-translation orbitals (32); 1 QP plane
  square8    8x8 square, t=1, U=4, 32+32, 2x2-sublattice orbitals (256),
             Gutzwiller (1), distance Jastrow; 8 QP planes (2 spin GL points
             x 4 translations)
  square16   16x16 square, t=1, t'=-0.3, U=8, 112+112 (1/8 hole doping),
             4x4-sublattice orbitals (4096), per-sublattice Gutzwiller (16),
             (sublattice, displacement) Jastrow; 8 QP planes
  uneven     (write_uneven_batch) 24 independent L x L square-lattice runs,
             L in {6, 8, 10, 12} (6 each), U ~ uniform[2, 10] and filling ~
             uniform[0.75, 1.0] from a seed (Nsize 28..144, equal spin
             counts), 2x2-sublattice orbitals, Gutzwiller, distance Jastrow,
             4 QP planes (4 translations), 2048 walkers each
Initial parameters are uniform[-0.1, 0.1] from a generator seed, written to
initial.def (six leading numbers, then (re, im, 0) per parameter).
"""
import os

import numpy as np

WORKLOADS = {
    "chain16": dict(kind="chain", L=16, t=1.0, tp=0.0, U=4.0, nup=8, sub=(2, 1), spgl=1, ntrans=1,
                    walkers=256, jastrow="nn"),
    "square8": dict(kind="square", L=8, t=1.0, tp=0.0, U=4.0, nup=32, sub=(2, 2), spgl=2, ntrans=4,
                    walkers=4096, jastrow="distance", gutz="single"),
    "square16": dict(kind="square", L=16, t=1.0, tp=-0.3, U=8.0, nup=112, sub=(4, 4), spgl=2, ntrans=4,
                     walkers=16384, jastrow="sublattice", gutz="sublattice"),
}

HEADER = "=============================================\n"

UNEVEN_SIDES = (6, 8, 10, 12)


def _sites(kind, L):
    if kind == "chain":
        return [(x, 0) for x in range(L)], (L, 1)
    return [(x, y) for y in range(L) for x in range(L)], (L, L)


def _idx(x, y, dims):
    Lx, Ly = dims
    return (x % Lx) + (y % Ly) * Lx


def write_workload(name, out_dir, seed=2024, sr_steps=10, nvmc_sample=1, interval=16, warmup=16,
                   rnd_seed=11272, step_dt=0.01, sta_del=1e-3, red_cut=1e-10, cg_max_iter=0, ex_update_path=0,
                   overrides=None):
    cfg = dict(WORKLOADS[name])
    if overrides:
        cfg.update(overrides)
    os.makedirs(out_dir, exist_ok=True)
    kind, L = cfg["kind"], cfg["L"]
    sites, dims = _sites(kind, L)
    ns = len(sites)
    Lx, Ly = dims
    sx, sy = cfg["sub"]
    nsub = sx * sy

    def sub(i):
        x, y = sites[i]
        return (x % sx) + (y % sy) * sx

    def disp(i, j):
        (xi, yi), (xj, yj) = sites[i], sites[j]
        return ((xj - xi) % Lx) + ((yj - yi) % Ly) * Lx

    # ---- Hamiltonian ----
    trans = []
    nn = [(1, 0)] if kind == "chain" else [(1, 0), (0, 1)]
    nnn = [] if (kind == "chain" or cfg["tp"] == 0.0) else [(1, 1), (1, -1)]
    for i, (x, y) in enumerate(sites):
        for vecs, t in ((nn, cfg["t"]), (nnn, cfg["tp"])):
            for dx, dy in vecs:
                j = _idx(x + dx, y + dy, dims)
                for s in (0, 1):
                    trans.append((j, s, i, s, t))
                    trans.append((i, s, j, s, t))
    with open(os.path.join(out_dir, "trans.def"), "w") as f:
        f.write("======================== \n")
        f.write(f"NTransfer {len(trans):6d}  \n")
        f.write("======================== \n========i_j_s_tijs====== \n======================== \n")
        for (i, s, j, t_, v) in trans:
            f.write(f"{i:5d} {s:5d} {j:5d} {t_:5d} {v:25.15f} {0.0:25.15f}\n")
    with open(os.path.join(out_dir, "coulombintra.def"), "w") as f:
        f.write(HEADER + f"NCoulombIntra {ns:10d}\n" + HEADER + "================== CoulombIntra ================\n" + HEADER)
        for i in range(ns):
            f.write(f"{i:5d} {cfg['U']:25.15f}\n")
    with open(os.path.join(out_dir, "locspn.def"), "w") as f:
        f.write("================================ \nNlocalSpin     0  \n================================ \n"
                "========i_1LocSpn_0IteElc ====== \n================================ \n")
        for i in range(ns):
            f.write(f"{i:5d} {0:6d}\n")

    # ---- Gutzwiller ----
    if cfg.get("gutz", "single") == "sublattice":
        gidx = [sub(i) for i in range(ns)]
        ng = nsub
    else:
        gidx = [0] * ns
        ng = 1
    with open(os.path.join(out_dir, "gutzwilleridx.def"), "w") as f:
        f.write(HEADER + f"NGutzwillerIdx {ng:10d}\nComplexType {0:10d}\n" + HEADER + HEADER)
        for i in range(ns):
            f.write(f"{i:5d} {gidx[i]:6d}\n")
        for k in range(ng):
            f.write(f"{k:5d} {1:6d}\n")

    # ---- Jastrow: one class per unordered pair, symmetric table ----
    classes = {}
    jflags = {}
    jmat = np.full((ns, ns), -1, dtype=np.int64)

    def jkey(i, j):
        a, b = min(i, j), max(i, j)
        if cfg["jastrow"] == "nn":
            d = min(disp(a, b), Lx - disp(a, b)) if kind == "chain" else None
            return ("nn",) if d == 1 else ("far",)
        if cfg["jastrow"] == "distance":
            (xa, ya), (xb, yb) = sites[a], sites[b]
            dx = min((xb - xa) % Lx, (xa - xb) % Lx)
            dy = min((yb - ya) % Ly, (ya - yb) % Ly)
            return (min(dx, dy), max(dx, dy))
        return (sub(a), disp(a, b))

    for i in range(ns):
        for j in range(ns):
            if i == j:
                continue
            k = jkey(i, j)
            if k not in classes:
                classes[k] = len(classes)
                jflags[classes[k]] = 0 if k == ("far",) else 1
            jmat[i, j] = classes[k]
    nj = len(classes)
    with open(os.path.join(out_dir, "jastrowidx.def"), "w") as f:
        f.write(HEADER + f"NJastrowIdx {nj:10d}\nComplexType {0:10d}\n" + HEADER + HEADER)
        for i in range(ns):
            for j in range(ns):
                if i != j:
                    f.write(f"{i:5d} {j:6d} {jmat[i, j]:6d}\n")
        for k in range(nj):
            f.write(f"{k:5d} {jflags[k]:6d}\n")

    # ---- orbitals: (sublattice of i, displacement j - i) ----
    omap = {}
    omat = np.zeros((ns, ns), dtype=np.int64)
    for i in range(ns):
        for j in range(ns):
            k = (sub(i), disp(i, j))
            if k not in omap:
                omap[k] = len(omap)
            omat[i, j] = omap[k]
    no = len(omap)
    with open(os.path.join(out_dir, "orbitalidx.def"), "w") as f:
        f.write(HEADER + f"NOrbitalIdx {no:10d}\nComplexType {0:10d}\n" + HEADER + HEADER)
        for i in range(ns):
            for j in range(ns):
                f.write(f"{i:5d} {j:6d} {omat[i, j]:6d}\n")
        for k in range(no):
            f.write(f"{k:5d} {1:6d}\n")

    # ---- translations ----
    shifts = [(0, 0)] if cfg["ntrans"] == 1 else [(0, 0), (1, 0), (0, 1), (1, 1)][:cfg["ntrans"]]
    if kind == "chain" and cfg["ntrans"] > 1:
        shifts = [(k, 0) for k in range(cfg["ntrans"])]
    with open(os.path.join(out_dir, "qptransidx.def"), "w") as f:
        f.write(HEADER + f"NQPTrans {len(shifts):10d}\n" + HEADER + "======== TrIdx_TrWeight_and_TrIdx_i_xi ======\n" + HEADER)
        for k in range(len(shifts)):
            f.write(f"{k} {1.0:10.5f}\n")
        for k, (dx, dy) in enumerate(shifts):
            for i, (x, y) in enumerate(sites):
                f.write(f"{k:5d} {i:6d} {_idx(x + dx, y + dy, dims):6d} {1:6d}\n")

    nelec_total = 2 * cfg["nup"]
    with open(os.path.join(out_dir, "modpara.def"), "w") as f:
        f.write("--------------------\nModel_Parameters   0\n--------------------\nVMC_Cal_Parameters\n"
                "--------------------\nCDataFileHead  zvo\nCParaFileHead  zqp\n--------------------\n")
        kv = [("NVMCCalMode", 0), ("NLanczosMode", 0), ("NDataIdxStart", 1), ("NDataQtySmp", 1),
              ("Nsite", ns), ("Ncond", nelec_total), ("2Sz", 0), ("NSPGaussLeg", cfg["spgl"]), ("NSPStot", 0),
              ("NMPTrans", len(shifts)), ("NSROptItrStep", sr_steps), ("NSROptItrSmp", sr_steps),
              ("DSROptRedCut", f"{red_cut:.10f}"), ("DSROptStaDel", f"{sta_del:.10f}"),
              ("DSROptStepDt", f"{step_dt:.10f}"), ("NSROptCGMaxIter", cg_max_iter),
              ("NVMCWarmUp", warmup), ("NVMCInterval", interval), ("NVMCSample", nvmc_sample),
              ("NExUpdatePath", ex_update_path), ("RndSeed", rnd_seed), ("NSplitSize", 1), ("NStore", 0), ("NSRCG", 1),
              ("NSRCGAbortOnFail", 0)]
        for k, v in kv:
            f.write(f"{k:<14s} {v}\n")
    with open(os.path.join(out_dir, "namelist.def"), "w") as f:
        f.write("ModPara modpara.def\nLocSpin locspn.def\nTrans trans.def\nCoulombIntra coulombintra.def\n"
                "Gutzwiller gutzwilleridx.def\nJastrow jastrowidx.def\nOrbital orbitalidx.def\n"
                "TransSym qptransidx.def\n")
    # ---- initial parameters uniform[-0.1, 0.1] ----
    rng = np.random.default_rng(seed)
    npara = ng + nj + no
    vals = rng.uniform(-0.1, 0.1, npara)
    for k in range(nj):  # frozen classes start (and stay) at zero
        if jflags[k] == 0:
            vals[ng + k] = 0.0
    with open(os.path.join(out_dir, "initial.def"), "w") as f:
        f.write(" ".join(["0.0"] * 6) + "\n")
        for v in vals:
            f.write(f"{v:.17e} 0.0 0.0\n")
    return {"namelist": os.path.join(out_dir, "namelist.def"), "initial": os.path.join(out_dir, "initial.def"),
            "nsite": ns, "nsize": nelec_total, "n_gutz": ng, "n_jast": nj, "n_slater": no, "npara": npara,
            "nqp": cfg["spgl"] * len(shifts), "walkers": cfg["walkers"], "ntransfer": len(trans)}


def write_uneven_batch(out_dir, seed=24, walkers=2048, runs_per_side=6, **kw):
    """The uneven batch: runs_per_side runs per side L in UNEVEN_SIDES, each a
    square L x L Hubbard model with t = 1, U ~ uniform[2, 10], filling ~
    uniform[0.75, 1.0] (electrons rounded to an even count with equal spins,
    at least 0.75 of the sites), 2x2-sublattice orbitals, Gutzwiller,
    distance Jastrow and 4 translation QP planes.  Every run gets its own
    directory run<k>/ and parameter seed (seed + k); keyword arguments go to
    write_workload (SR steps, warm-up, ...).  Returns the runs' infos, each
    with its "U" and "filling"."""
    rng = np.random.default_rng(seed)
    infos = []
    k = 0
    for L in UNEVEN_SIDES:
        for _ in range(runs_per_side):
            ns = L * L
            U = float(rng.uniform(2.0, 10.0))
            f = float(rng.uniform(0.75, 1.0))
            nup = int(round(f * ns / 2))
            nup = max(-(-3 * ns // 8), min(ns // 2, nup))
            ov = dict(kind="square", L=L, t=1.0, tp=0.0, U=U, nup=nup, sub=(2, 2), spgl=1, ntrans=4,
                      walkers=walkers, jastrow="distance", gutz="single")
            info = write_workload("square8", os.path.join(out_dir, f"run{k}"), seed=seed + k, overrides=ov, **kw)
            info.update(U=U, filling=2 * nup / ns, L=L)
            infos.append(info)
            k += 1
    return infos
