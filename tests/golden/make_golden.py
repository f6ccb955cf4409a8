"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

  tebd_n64_d20_chi256.npz   the TEBD brickwork workload (n=64, d=2, |0...0>,
                            depth 20, Haar gates seeded 20261202+1, serpentine
                            order, maxdim 256, absolute threshold 1e-10):
                            1024 seeded bitstrings and their FP64 oracle
                            amplitudes, the final bond dimensions, and the
                            nominal flop counts of the sweep.
  tebd_vidal_n64_d20_chi256.npz   the same gates on the same bonds (each
                            brickwork layer's gates together) in the Vidal
                            gauge (oracle.vidal_gate): the same 1024
                            bitstrings' amplitudes, bond dimensions and the
                            Lambda spectrum of the middle bond (--vidal).

Run:  OPENBLAS_NUM_THREADS=8 python tests/golden/make_golden.py [--tag T]
A second run with another thread count gives the oracle's own
reproducibility (stored as ``self_relerr`` when --compare is passed).
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import tenet_oracle as O  # noqa: E402

N, DEPTH, CHI, THR = 64, 20, 256, 1e-10
SEED = O.BASE_SEED + 1
NAMP = 1024


def bitstrings(n, count, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 2, size=(count, n), dtype=np.uint8)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tebd_n64_d20_chi256.npz"))
    ap.add_argument("--compare", default=None, help="an earlier fixture to compare against")
    ap.add_argument("--vidal", action="store_true", help="the Vidal-gauge fixture")
    args = ap.parse_args()
    sched, gates = O.tebd_gates(SEED, N, DEPTH)
    if args.vidal:
        return vidal(args, sched, gates)
    t0 = time.perf_counter()
    last = [t0]

    def progress(g):
        if g % 63 == 62:
            now = time.perf_counter()
            print(f"layer pair {g // 63 + 1}/10  {now - last[0]:.1f}s", flush=True)
            last[0] = now

    psi = O.tebd_run(N, sched, gates, CHI, THR, progress)
    dt = time.perf_counter() - t0
    bits = bitstrings(N, NAMP, SEED + 7)
    amps = np.array([O.amplitude(b, psi) for b in bits])
    out = dict(bits=np.packbits(bits, axis=1), amps=amps, bonds=np.array(psi.bondsizes()), seconds=dt,
               threads=os.environ.get("OPENBLAS_NUM_THREADS", "default"))
    if args.compare:
        ref = np.load(args.compare)
        rel = np.linalg.norm(amps - ref["amps"]) / np.linalg.norm(ref["amps"])
        out["self_relerr"] = rel
        print("relative l2 difference to", args.compare, ":", rel)
    np.savez(args.out, **out)
    print(f"sweep {dt:.1f}s, max bond {max(psi.bondsizes())}, wrote {args.out}")


def layers_of(sched, n):
    """Split the serpentine schedule into its brickwork layers (order kept)."""
    out, i, parity = [], 0, 0
    while i < len(sched):
        cnt = len(range(parity, n - 1, 2))
        out.append(list(sched[i:i + cnt]))
        i += cnt
        parity ^= 1
    return out


def vidal(args, sched, gates):
    out_path = args.out if args.out != os.path.join(HERE, "tebd_n64_d20_chi256.npz") else \
        os.path.join(HERE, "tebd_vidal_n64_d20_chi256.npz")
    layers = layers_of(sched, N)
    t0 = time.perf_counter()
    v = O.vidal_tebd_run(N, layers, gates, CHI, THR)
    dt = time.perf_counter() - t0
    psi = v.to_mps()
    bits = bitstrings(N, NAMP, SEED + 7)
    amps = np.array([O.amplitude(b, psi) for b in bits])
    np.savez(out_path, bits=np.packbits(bits, axis=1), amps=amps, bonds=np.array(v.bondsizes()),
             lam_mid=v.lam[N // 2 - 1], seconds=dt, threads=os.environ.get("OPENBLAS_NUM_THREADS", "default"))
    print(f"vidal sweep {dt:.1f}s, max bond {max(v.bondsizes())}, wrote {out_path}")


if __name__ == "__main__":
    main()
