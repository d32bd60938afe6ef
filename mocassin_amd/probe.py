"""The device's rate of random-address atomic adds (mc_probe_atomics), the
ceiling of the transport's J updates: against the array's footprint (FP64,
64 active lanes), against the active lanes per wavefront (FP64 at 583 MB,
the 45^3 workload's J3), and FP32 at 583 MB.

    python -m mocassin_amd.probe > probe.json
"""
from .transport import probe_atomics


def sweep(requests=1 << 30, log=print):
    out = []

    def one(mb, lanes, fp64, req):
        words = mb * (1 << 20) // (8 if fp64 else 4)
        ms, rate = probe_atomics(words, req, lanes=lanes, fp64=fp64)
        out.append({"footprint_MB": mb, "lanes": lanes, "fp64": fp64, "requests": req, "kernel_ms": ms,
                    "adds_per_s": rate})
        log("%s %5d MB, %2d lanes: %.2f ms, %.3g adds/s" % ("FP64" if fp64 else "FP32", mb, lanes, ms, rate))
    for mb in (4096, 583, 292, 146, 73, 32, 16, 4, 1):
        one(mb, 64, True, requests)
    for lanes in (1, 4, 16, 32, 64):
        one(583, lanes, True, requests // 4)
    one(583, 64, False, requests)
    return out


if __name__ == "__main__":
    import json
    import sys
    print(json.dumps(sweep(log=lambda m: print(m, file=sys.stderr, flush=True))))
