"""The device's rate of random-address atomic adds (mc_probe_atomics), the
ceiling of the transport's J updates: against the array's footprint (FP64,
64 active lanes), against the active lanes per wavefront (FP64 at 583 MB,
the 45^3 workload's J3), and FP32 at 583 MB.  Then the same adds with the
lanes of each instruction sharing 64-byte lines (mc_probe_atomic_lines):
sets of 1, 2, 4 or 8 lanes per line, adjacent or spread across the
wavefront, at 583 MB.

    python -m mocassin_amd.probe > probe.json
"""
from .transport import probe_atomic_lines, probe_atomics


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
    words = 583 * (1 << 20) // 8
    for group, spread in ((1, False), (2, False), (4, False), (8, False), (2, True), (8, True)):
        ms, rate, lines = probe_atomic_lines(words, requests, group, spread)
        out.append({"footprint_MB": 583, "lanes_per_line": group, "spread": spread, "requests": requests,
                    "kernel_ms": ms, "adds_per_s": rate, "line_requests_per_s": lines})
        log("FP64 583 MB, %d lanes per 64-byte line (%s): %.2f ms, %.3g adds/s, %.3g lines/s"
            % (group, "spread" if spread else "adjacent", ms, rate, lines))
    return out


if __name__ == "__main__":
    import json
    import sys
    print(json.dumps(sweep(log=lambda m: print(m, file=sys.stderr, flush=True))))
