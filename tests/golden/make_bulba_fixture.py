"""Cuts the first 256 Gaussians of the reference's 3DGS asset
(asset/bulba.ply, 11982 Gaussians in the 62-float layout) into
tests/golden/bulba_first256.ply: the same header with the vertex count set
to 256, then the first 256 records unchanged. Run once in a checkout that
has the reference next to it; the output is committed."""
import os
import sys

SRC = "/root/reference/asset/bulba.ply"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bulba_first256.ply")
N = 256


def main(src=SRC):
    with open(src, "rb") as f:
        lines = []
        while True:
            line = f.readline()
            lines.append(line)
            if line == b"end_header\n":
                break
        records = f.read(62 * 4 * N)
    out = []
    for line in lines:
        if line.startswith(b"element vertex "):
            line = b"element vertex %d\n" % N
        out.append(line)
    assert len(records) == 62 * 4 * N
    with open(OUT, "wb") as f:
        f.write(b"".join(out) + records)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main(*sys.argv[1:])
