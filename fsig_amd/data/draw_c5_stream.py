"""Draws the C5 request stream on the device (each candidate query is checked
against the stream's bounds) and writes it as JSON: python -m
fsig_amd.data.draw_c5_stream OUT.json.  The bench reads
fsig_amd/data/c5_stream.json when its parameters match."""
import sys
import time

import fsig_amd as F
from fsig_amd import workloads as W


def main(out):
    t = time.time()
    g = W.c5_graph()
    print(f"graph {g.n} V in {time.time() - t:.1f}s", flush=True)
    ctx = F.Context(0, 32 << 30)
    dg = ctx.upload(g)
    W.c5_stream(g, dg, progress=lambda m: print(m, flush=True), cache=out)
    print("written", out, flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else W.C5_STREAM_FILE)
