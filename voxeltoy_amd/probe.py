"""Device diagnostics.

  python -m voxeltoy_amd.probe copy [--bytes N] [--iters K]
      streaming copy of N bytes (read N + write N); prints the kernel time and
      the achieved HBM bandwidth.  Used as the known-size reference for the
      FETCH_SIZE / WRITE_SIZE hardware counters.
"""
import argparse
import json

from . import api


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["copy"])
    ap.add_argument("--bytes", type=int, default=1 << 30)
    ap.add_argument("--iters", type=int, default=10)
    a = ap.parse_args()
    with api.Context(0) as ctx:
        ms = ctx.copy_probe(a.bytes, a.iters)
    print(json.dumps({"probe": "copy", "bytes": a.bytes, "kernel_ms": ms,
                      "gbps_read_plus_write": 2 * a.bytes / (ms / 1e3) / 1e9}))


if __name__ == "__main__":
    main()
