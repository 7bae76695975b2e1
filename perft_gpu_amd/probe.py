"""Measure this device's ceilings with the engine's probe kernels.

    python -m perft_gpu_amd.probe                 # all probes, JSON line
    python -m perft_gpu_amd.probe --kind copy --bytes 268435456 --iters 1

kinds: valu (int32 lane-ops/s), copy (read+write bytes/s), random16 (16-byte
read-modify-writes/s over a --bytes table), opcodes (lane-ops/s of single
instruction kinds in 8 independent chains)."""
import argparse
import json

from . import Engine

KINDS = {"valu": 0, "copy": 1, "random16": 2, "opcodes": 3}
OPCODES = ["v_add_u32", "v_xor_b32", "v_and_b32", "v_bcnt_u32_b32", "v_lshlrev_b32", "v_mul_lo_u32", "v_lshlrev_b64",
           "v_bitop3_b32", "v_and_or_b32", "v_or3_b32", "v_cndmask_b32", "v_perm_b32", "v_add3_u32", "v_min_u32",
           "v_ashrrev_i32", "v_bfe_i32"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--kind", choices=list(KINDS) + ["all"], default="all")
    ap.add_argument("--bytes", type=int, default=0)
    ap.add_argument("--iters", type=int, default=0)
    a = ap.parse_args()
    eng = Engine(arena_bytes=64 << 20)
    out = {}
    kinds = list(KINDS) if a.kind == "all" else [a.kind]
    for k in kinds:
        if k == "random16" and a.kind == "all":
            for gib in (1, 16, 64):
                out["random16_%dgib_per_s" % gib] = eng.probe(2, gib << 30, a.iters)
        elif k == "opcodes":
            for i, name in enumerate(OPCODES):
                out["lane_ops_per_s_" + name] = eng.probe(3, 0, i)
        else:
            out[k + "_per_s"] = eng.probe(KINDS[k], a.bytes, a.iters)
    eng.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
