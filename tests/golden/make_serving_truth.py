"""Reference permeabilities of the serving benchmark's request stream (needs a GPU).

The stream is the one bench.py's serving leg draws: seed 7, 32 requests, edges uniform in
48..160 lu, thresholded Gaussian random fields (correlation length 6 lu, porosity uniform in
0.15..0.45), nu = 1/6, g = (1e-6, 0, 0), k along x.  Each k is an FP64 solve run far past the
benchmark's accuracy target: the FP64 server with Anderson acceleration until k changes by
<= 1e-11 relative between accelerations (twice), cross-checked against plain FP64 stepping.

    python3 tests/golden/make_serving_truth.py   # writes tests/golden/serving_stream_k.json
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "tests"))
from conftest import load_package  # noqa: E402

SEED, NREQ = 7, 32


def stream(pkg):
    samples, _, edges, pors = pkg.request_stream(NREQ, 200.0, SEED)
    return [(int(e), float(p), s) for e, p, s in zip(edges, pors, samples)]


def main():
    pkg = load_package()
    reqs = stream(pkg)
    with pkg.PermeabilityServer(check_interval=100, tol=1e-11, max_steps=400000, block=256, precision="fp64",
                                accel_depth=5, accel_period=200, accel_warmup=500) as srv:
        tickets = [srv.submit(s, 1.0 / 6.0, (1e-6, 0.0, 0.0)) for (_, _, s) in reqs]
        srv.wait_idle()
        res = {r["ticket"]: r for r in srv.poll()}
    out = []
    for i, t in enumerate(tickets):
        r = res[t]
        out.append({"index": i, "edge": reqs[i][0], "porosity": reqs[i][1], "fluid_nodes": int(r["fluid_nodes"]),
                    "k": float(r["k"][0]), "steps": int(r["steps"]), "converged": int(r["converged"]),
                    "last_rel_change": float(r["rel_change"])})
        print(f"{i:2d} edge {reqs[i][0]:3d} porosity {reqs[i][1]:.3f} k {r['k'][0]:.12e} steps {r['steps']} "
              f"conv {r['converged']} rel {r['rel_change']:.1e}", flush=True)
    # plain FP64 cross-check on the first four samples
    for i in range(4):
        s = pkg.SinglePhase(reqs[i][2], 1.0 / 6.0, (1e-6, 0.0, 0.0), precision="fp64")
        s.step(60000)
        k = s.Comput_Perm(1)[0][0]
        out[i]["plain_fp64_60000_steps"] = float(k)
        print(f"{i:2d} plain FP64 after 60000 steps {k:.12e} (rel diff {abs(k - out[i]['k']) / out[i]['k']:.1e})",
              flush=True)
        s.close()
    path = os.path.join(ROOT, "tests", "golden", "serving_stream_k.json")
    with open(path, "w") as f:
        json.dump({"seed": SEED, "requests": NREQ, "nu": 1.0 / 6.0, "g": [1e-6, 0.0, 0.0], "samples": out}, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
