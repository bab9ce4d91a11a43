"""PLY files for the loader tests: the header shape the reference's readers
assert (io_ply.rs:30-105, :128-227) and little-endian records."""
import numpy as np

GAUSS_PROPS = (["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
               + [f"f_rest_{i}" for i in range(45)]
               + ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"])


def write_gauss(path, records, comment="test file"):
    """records: float32 [N, 62] in the order of GAUSS_PROPS."""
    records = np.ascontiguousarray(records, "<f4")
    head = ["ply", "format binary_little_endian 1.0", f"comment {comment}",
            f"element vertex {len(records)}"]
    head += [f"property float {p}" for p in GAUSS_PROPS] + ["end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        f.write(records.tobytes())


def write_xyzrgb(path, xyz, rgb):
    dt = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("r", "u1"), ("g", "u1"),
                   ("b", "u1")])
    a = np.zeros(len(xyz), dt)
    a["x"], a["y"], a["z"] = np.asarray(xyz, np.float64).T
    a["r"], a["g"], a["b"] = np.asarray(rgb, np.uint8).T
    head = ["ply", "format binary_little_endian 1.0", "comment test", f"element vertex {len(a)}",
            "property double x", "property double y", "property double z",
            "property uchar red", "property uchar green", "property uchar blue", "end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        f.write(a.tobytes())
