"""Cuts the reference's 3DGS asset (asset/bulba.ply, 11982 Gaussians in the
62-float layout) into two fixtures under tests/golden/, each the same header
with the vertex count changed, then records copied unchanged:

* bulba_first256.ply: the first 256 records;
* bulba_every2nd.ply: records 0, 2, 4, ... (5991 Gaussians spread over the
  whole object), the real-data scene of the full-chain GPU parity test.

Run once in a checkout that has the reference next to it; the outputs are
committed."""
import os
import sys

SRC = "/root/reference/asset/bulba.ply"
HERE = os.path.dirname(os.path.abspath(__file__))
REC = 62 * 4


def _read(src):
    with open(src, "rb") as f:
        lines = []
        while True:
            line = f.readline()
            lines.append(line)
            if line == b"end_header\n":
                break
        body = f.read()
    num = next(int(l.split()[2]) for l in lines if l.startswith(b"element vertex "))
    assert len(body) == REC * num
    return lines, [body[i * REC:(i + 1) * REC] for i in range(num)]


def _write(path, lines, records):
    out = [b"element vertex %d\n" % len(records) if l.startswith(b"element vertex ") else l
           for l in lines]
    with open(path, "wb") as f:
        f.write(b"".join(out) + b"".join(records))
    print(f"wrote {path} ({len(records)} Gaussians)")


def main(src=SRC):
    lines, recs = _read(src)
    _write(os.path.join(HERE, "bulba_first256.ply"), lines, recs[:256])
    _write(os.path.join(HERE, "bulba_every2nd.ply"), lines, recs[::2])


if __name__ == "__main__":
    main(*sys.argv[1:])
