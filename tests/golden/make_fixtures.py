#!/usr/bin/env python3
"""Write the committed golden fixtures of tests/golden from the reference
checkout's own fixture data (run once where the reference checkout is
mounted; the tests only read the files written here).

Everything written here is data held by the reference's tests:
  * SFMT words of the C library      crates/sfmt19937/tests/golden_vs_c.rs:40-160
  * Hubbard chain step-1 operands    tests/fixtures/c_order_sr_operands/hubbard-cg-store0.txt
    and its Expert inputs             tests/fixtures/c_order_sr_operands/inputs/hubbard
    (modpara overrides as in tests/fixtures/c_order_sr_operands/PROVENANCE.md)
  * Heisenberg multi-rank state      tests/fixtures/mpi_matrix_179/real-{cg,d0}-r{2,4}w1-s7, real-d1-r2w1-s7
    and its Expert inputs             tests/fixtures/native_c_physcal_181/heisenberg_chain_real/inputs
    (overrides from tests/fixtures/mpi_matrix_179/cells.txt)
  * direct-SR solutions (dposv)      tests/fixtures/negative_stepdt/c_sr.txt
  * SR-CG iterates                   tests/fixtures/sr_cg/c_refresh/real{,_diag_scale}.txt
  * PfaPack input matrices           tests/fixtures/pfapack/real_n*.orig.txt
                                     tests/fixtures/pfapack/complex_n*.orig.txt
  * two-hop bilinear bits            tests/fixtures/pfaffian_cg/two_hop_bilinear{,_c}.txt
  * hopping candidate retry tapes    tests/fixtures/candidate/hopping_retry_cap_559.txt
  * zvo_out row layout               tests/fixtures/c_integration_reference/hubbard_chain_real/zvo_out_first10.dat
  * ten C optimisation steps         tests/fixtures/c_integration_reference/heisenberg_chain_real/zvo_out_first10.dat
  * C SetFlagShift gates             tests/fixtures/optimization_flags/c_shift_flags.txt

Usage: make_fixtures.py [--ref REFERENCE_ROOT] [fixture ...] (default: all).
"""
import json
import os
import re
import shutil
import struct
import sys

import numpy as np

ARGS = sys.argv[1:]
REF = "/root/reference"
if len(ARGS) >= 2 and ARGS[0] == "--ref":
    REF, ARGS = ARGS[1], ARGS[2:]
OUT = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(REF, "tests", "fixtures")


def hexd(s):
    return struct.unpack(">d", bytes.fromhex(s))[0]


def sfmt():
    src = open(os.path.join(REF, "crates/sfmt19937/tests/golden_vs_c.rs")).read()
    first8 = {}
    block = src[src.index("GOLDEN_FIRST8_U32"):src.index("GOLDEN_BOUNDARY_SEED")]
    for seed, body in re.findall(r"\(\s*([0-9_]+|u32::MAX),\s*\[([^\]]*)\]", block):
        key = 0xFFFFFFFF if seed == "u32::MAX" else int(seed.replace("_", ""))
        first8[str(key)] = [int(x.replace("_", ""), 16) for x in re.findall(r"0x[0-9a-f_]+", body)]
    bblock = src[src.index("const GOLDEN_BOUNDARY:"):src.index("GOLDEN_FIRST8_U64_SEED")]
    boundary = [[int(i), int(v.replace("_", ""), 16)] for i, v in re.findall(r"\((\d+),\s*(0x[0-9a-f_]+)\)", bblock)]
    u64block = src[src.index("const GOLDEN_FIRST8_U64:"):]
    u64block = u64block[:u64block.index("];")]
    u64 = [int(v.replace("_", ""), 16) for v in re.findall(r"0x[0-9a-f_]+", u64block)]
    json.dump({"first8_u32": first8, "boundary_seed": 11272, "boundary": boundary,
               "first8_u64_seed": 11272, "first8_u64": u64},
              open(os.path.join(OUT, "sfmt_golden.json"), "w"), indent=1)


def copy_inputs(src_dir, dst_dir, overrides):
    os.makedirs(dst_dir, exist_ok=True)
    for name in os.listdir(src_dir):
        shutil.copy(os.path.join(src_dir, name), os.path.join(dst_dir, name))
        os.chmod(os.path.join(dst_dir, name), 0o644)
    path = os.path.join(dst_dir, "modpara.def")
    lines = open(path).read().split("\n")
    seen = set()
    out = []
    for line in lines:
        t = line.split()
        if t and t[0] in overrides:
            out.append(f"{t[0]} {overrides[t[0]]}")
            seen.add(t[0])
        else:
            out.append(line)
    while out and out[-1] == "":
        out.pop()
    for k, v in overrides.items():
        if k not in seen:
            out.append(f"{k} {v}")
    open(path, "w").write("\n".join(out) + "\n")


def hubbard():
    copy_inputs(os.path.join(FIX, "c_order_sr_operands/inputs/hubbard"), os.path.join(OUT, "inputs/hubbard_chain"),
                {"RndSeed": 1, "NSROptItrStep": 1, "NSROptItrSmp": 1, "NSRCG": 1, "NStore": 0,
                 "NSRCGAbortOnFail": 0})
    ref = {}
    for line in open(os.path.join(FIX, "c_order_sr_operands/hubbard-cg-store0.txt")):
        if line.startswith("#") or not line.strip():
            continue
        t = line.split()
        ref[t[0]] = t[1:]
    json.dump(ref, open(os.path.join(OUT, "hubbard_cg_store0.json"), "w"), indent=0)


def heisenberg():
    over = {"NVMCCalMode": 0, "NSROptItrStep": 1, "NSROptItrSmp": 1, "NVMCSample": 7, "NSplitSize": 1,
            "NStore": 0, "NSRCG": 1, "NSRCGAbortOnFail": 0}
    copy_inputs(os.path.join(FIX, "native_c_physcal_181/heisenberg_chain_real/inputs"),
                os.path.join(OUT, "inputs/heisenberg_chain"), over)
    for cell in ["real-cg-r2w1-s7", "real-cg-r4w1-s7", "real-d0-r2w1-s7", "real-d0-r4w1-s7", "real-d1-r2w1-s7"]:
        base = os.path.join(FIX, "mpi_matrix_179", cell)
        ranks = []
        r = 0
        while os.path.exists(os.path.join(base, f"rank{r}.txt")):
            rec = {"eleidx": {}, "counter": None, "rng": None, "rng_idx": None}
            for line in open(os.path.join(base, f"rank{r}.txt")):
                t = line.split()
                if not t:
                    continue
                if t[0] == "COUNTER":
                    rec["counter"] = [int(x) for x in t[3:]]
                elif t[0] == "ELEIDX":
                    rec["eleidx"][t[3].split("=")[1]] = [int(x) for x in t[4:]]
                elif t[0] == "RNG":
                    rec["rng_idx"] = int(t[3].split("=")[1])
                    rec["rng"] = [int(x) for x in t[4:]]
            ranks.append(rec)
            r += 1
        ops = {}
        for line in open(os.path.join(base, "operands.txt")):
            t = line.split()
            if t:
                ops[t[0]] = [float(x) for x in t[1:]]
        json.dump({"ranks": ranks, "operands": ops}, open(os.path.join(OUT, f"mpi179_{cell.replace('-', '_')}.json"), "w"))


def negative_stepdt():
    """C 1.4.0 direct-SR solutions (stcOptInit + dposv) on fixed operands for
    positive, negative and shifted DSROptStepDt (data lines, as JSON)."""
    cases = []
    for line in open(os.path.join(FIX, "negative_stepdt", "c_sr.txt")):
        if line.startswith("#") or not line.strip():
            continue
        t = line.split()
        if t[0] == "case":
            cases.append({"name": t[1], "raw_dt": t[3], "sta_del": t[5]})
        elif t[0] == "solution":
            cases[-1]["solution"] = [float(x) for x in t[1:]]
        elif t[0] in ("srflag", "info"):
            cases[-1][t[0]] = int(t[1])
        elif t[0] == "dt":
            cases[-1]["dt"] = float(t[1])
    json.dump(cases, open(os.path.join(OUT, "negative_stepdt_c_sr.json"), "w"), indent=0)


def sr_cg():
    for name in ["real.txt", "real_diag_scale.txt"]:
        shutil.copy(os.path.join(FIX, "sr_cg/c_refresh", name), os.path.join(OUT, "sr_cg_" + name))
        os.chmod(os.path.join(OUT, "sr_cg_" + name), 0o644)


def pfapack():
    arrays = {}
    for name in sorted(os.listdir(os.path.join(FIX, "pfapack"))):
        if not name.startswith("real_"):
            continue
        n = int(re.search(r"_n(\d+)_", name).group(1))
        if n > 128:
            continue
        vals = []
        for line in open(os.path.join(FIX, "pfapack", name)):
            line = line.strip()
            if not line or line.startswith("#") or " " in line:
                continue
            vals.append(float(line))
        arrays[name.replace(".orig.txt", "")] = np.array(vals, dtype=np.float64).reshape(n, n, order="F")
    np.savez_compressed(os.path.join(OUT, "pfapack_real_inputs.npz"), **arrays)


def pfapack_complex():
    """Complex skew-symmetric inputs (n <= 128) of the reference's identity
    tests, as complex128 column-major matrices (value = re + i im per line)."""
    arrays = {}
    for name in sorted(os.listdir(os.path.join(FIX, "pfapack"))):
        if not name.startswith("complex_"):
            continue
        n = int(re.search(r"_n(\d+)_", name).group(1))
        if n > 128:
            continue
        vals = []
        for line in open(os.path.join(FIX, "pfapack", name)):
            parts = line.split()
            if len(parts) != 2 or line.startswith("#") or not parts[0][0] in "+-0123456789":
                continue
            try:
                vals.append(complex(float(parts[0]), float(parts[1])))
            except ValueError:
                continue
        arrays[name.replace(".orig.txt", "")] = np.array(vals, dtype=np.complex128).reshape(n, n, order="F")
    np.savez_compressed(os.path.join(OUT, "pfapack_complex_inputs.npz"), **arrays)


def candidate():
    """C retry-cap tapes of the hopping candidate (data file, copied as is)."""
    shutil.copy(os.path.join(FIX, "candidate", "hopping_retry_cap_559.txt"),
                os.path.join(OUT, "candidate_hopping_retry_cap.txt"))
    os.chmod(os.path.join(OUT, "candidate_hopping_retry_cap.txt"), 0o644)


def zvo_rows():
    """C 1.4.0 zvo_out.dat rows of the real Hubbard chain (data file, as is):
    the byte layout of the per-step output row."""
    shutil.copy(os.path.join(FIX, "c_integration_reference", "hubbard_chain_real", "zvo_out_first10.dat"),
                os.path.join(OUT, "zvo_out_c_rows.dat"))
    os.chmod(os.path.join(OUT, "zvo_out_c_rows.dat"), 0o644)


def heisenberg_opt_rows():
    """C 1.4.0 zvo_out.dat rows of ten optimisation steps of the Heisenberg
    chain (c_integration_reference/heisenberg_chain_real, seed 1, direct SR,
    NStore 1, 100 samples; data file, as is) and the same model's Expert
    inputs in optimisation mode (NVMCCalMode 0; every other key as given)."""
    copy_inputs(os.path.join(FIX, "native_c_physcal_181/heisenberg_chain_real/inputs"),
                os.path.join(OUT, "inputs/heisenberg_chain_opt"), {"NVMCCalMode": 0})
    shutil.copy(os.path.join(FIX, "c_integration_reference", "heisenberg_chain_real", "zvo_out_first10.dat"),
                os.path.join(OUT, "zvo_out_heisenberg_real_first10.dat"))
    os.chmod(os.path.join(OUT, "zvo_out_heisenberg_real_first10.dat"), 0o644)


def shift_flags():
    """C SetFlagShift eligibility records (data file, copied as is)."""
    shutil.copy(os.path.join(FIX, "optimization_flags", "c_shift_flags.txt"), os.path.join(OUT, "c_shift_flags.txt"))
    os.chmod(os.path.join(OUT, "c_shift_flags.txt"), 0o644)


def bilinear():
    lines = [l for l in open(os.path.join(FIX, "pfaffian_cg/two_hop_bilinear.txt")) if l.strip() and not l.startswith("#")]
    expected = [l.split() for l in open(os.path.join(FIX, "pfaffian_cg/two_hop_bilinear_c.txt")) if not l.startswith("#")]
    sizes, values, bits = [], [], []
    it = iter(lines)
    for n_s, b in expected:
        n = int(n_s)
        assert int(next(it)) == n
        vals = [struct.unpack(">d", bytes.fromhex(v))[0] for v in next(it).split()]
        sizes.append(n)
        values.append(np.array(vals[1:], dtype=np.float64))
        bits.append(int(b, 16))
    np.savez_compressed(os.path.join(OUT, "two_hop_bilinear.npz"), sizes=np.array(sizes),
                        values=np.concatenate(values), bits=np.array(bits, dtype=np.uint64))


ALL = [sfmt, hubbard, heisenberg, sr_cg, pfapack, pfapack_complex, candidate, zvo_rows, heisenberg_opt_rows, bilinear,
       shift_flags, negative_stepdt]

if __name__ == "__main__":
    chosen = [f for f in ALL if not ARGS or f.__name__ in ARGS]
    for f in chosen:
        f()
    print("fixtures written to", OUT, ":", ", ".join(f.__name__ for f in chosen))
