"""Write tests/golden/config_hashes.json: the oracle's wave-synchronous
cycles on the two large workloads, reduced to SHA-256 digests of each field.

* the time-to-steady workload: 2048x2048 fBm seed 7, default parameters,
  262144 drops per cycle, cycles 0 and 1;
* the batched ensemble: the 64 terrains of ensemble_spec(64, 1) (sides
  128..1024, per-terrain evapRate/depositionRate, 1 drop per 16 cells, each
  terrain spawning from its own seed and terrain number), cycles 0 and 1.

The oracle needs about a minute per 2048x2048 cycle on one core, so the GPU
tests compare the device's fields with these digests instead of re-running it.
Digests are of the float32 arrays' bytes (x-major, little endian); each cycle
also records the oracle's step statistics.  Run from the repository root:
    python tests/golden/make_config_hashes.py [--jobs 8]
"""
import argparse
import hashlib
import json
import os
import sys
from multiprocessing import Pool

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config_hashes.json")
FIELDS = ("height", "discharge", "momentumx", "momentumy", "erf_discharge")
CYCLES = 2


def digest(a) -> str:
    import numpy as np
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


def run(job):
    import numpy as np

    import oracle
    name, t, nx, ny, drops, seed, kw = job
    h = oracle.fbm(seed, nx, ny)
    z = np.zeros_like(h)
    d, mx, my, root, erfd = z.copy(), z.copy(), z.copy(), z.copy(), z.copy()
    P = oracle.Params.default(**kw)
    out = {"nx": nx, "ny": ny, "drops": drops, "seed": seed, "terrain": t, "params": kw,
           "initial_height": digest(h), "cycles": []}
    for c in range(CYCLES):
        st = oracle.wave_cycle(h, d, mx, my, root, erfd, P, seed, c, drops, terrain=t)
        out["cycles"].append({"stats": st.as_dict(),
                              **{f: digest(a) for f, a in zip(FIELDS, (h, d, mx, my, erfd))}})
    print(name, t, nx, drops, "done", flush=True)
    return name, t, out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    import oracle
    import simplehydrology_amd as shy
    oracle.build()
    jobs = [("steady2048", 0, 2048, 2048, 262144, 7, {})]
    for t, s in enumerate(shy.ensemble_spec(64, 1)):
        jobs.append(("ensemble64", t, s["nx"], s["ny"], s["drops"], s["seed"],
                     {"evapRate": s["evapRate"], "depositionRate": s["depositionRate"]}))
    jobs.sort(key=lambda j: -j[2] * j[3])
    res = {"steady2048": None, "ensemble64": [None] * 64}
    with Pool(args.jobs) as pool:
        for name, t, out in pool.imap_unordered(run, jobs):
            if name == "steady2048":
                res[name] = out
            else:
                res[name][t] = out
    res["fields"] = list(FIELDS)
    res["made_by"] = "tests/golden/make_config_hashes.py (oracle wave cycle, oracle/shy_oracle.c)"
    with open(OUT, "w") as f:
        json.dump(res, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
