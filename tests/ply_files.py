"""PLY and .splat test inputs for the model-file tests (io/ply.rs), written
here byte by byte so that the files do not come from the code under test.

Each crafted file comes with the Gaussian3D records that load_ply must
produce, derived from io/ply.rs:116-247 by hand: Gaussian3D::new's defaults
(math/mod.rs:52-62: rotation (1, 0, 0, 0), log_scale -3, opacity logit -2,
everything else 0), then the properties in header order.
"""
import numpy as np

f32 = np.float32
# Gaussian3D record offsets
POS, OPL, ROT, LS, SH = 0, 3, 4, 8, 11


def default_record():
    g = np.zeros(59, f32)
    g[ROT] = 1.0
    g[LS:LS + 3] = -3.0
    g[OPL] = -2.0
    return g


def random_scene(n, seed):
    rng = np.random.default_rng(seed)
    g = rng.normal(0, 1, (n, 59)).astype(f32)
    g[:, LS:LS + 3] = rng.normal(-3, 1, (n, 3)).astype(f32)
    return g


def special_scene():
    """Records with values the rounding and clamping rules treat specially."""
    rows = []
    for v in [0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, 3.4e38, -3.4e38, 0.5, -0.5, 1.0, 10.0]:
        g = default_record()
        g[POS:POS + 3] = [v, 1.0, -1.0]
        g[OPL] = v
        g[SH] = v
        g[SH + 16] = -v if v == v else v
        g[SH + 32] = 0.5 * v if v == v else v
        g[ROT:ROT + 4] = [v, 0.0, 0.0, 0.0]
        g[LS:LS + 3] = [v if abs(v) < 80 else 80.0, -1.0, 0.0]
        rows.append(g)
    # the quaternion's zero/near-zero branch and a rounding-half case
    g = default_record()
    g[ROT:ROT + 4] = 0.0
    rows.append(g)
    g = default_record()
    g[ROT:ROT + 4] = [0.5, 0.5, 0.5, 0.5]
    g[SH] = (0.5 / 255.0 - 0.5) / 0.2820948
    rows.append(g)
    return np.stack(rows).astype(f32)


def reference_header(n):
    """The header text save_ply writes (io/ply.rs:43-64)."""
    lines = ["ply", "format binary_little_endian 1.0", f"element vertex {n}"]
    lines += [f"property float {p}" for p in ["x", "y", "z", "nx", "ny", "nz"]]
    lines += [f"property float f_dc_{i}" for i in range(3)]
    lines += [f"property float f_rest_{i}" for i in range(45)]
    lines += ["property float opacity"]
    lines += [f"property float scale_{i}" for i in range(3)]
    lines += [f"property float rot_{i}" for i in range(4)]
    lines += ["end_header"]
    return ("\n".join(lines) + "\n").encode()


def reference_body(g):
    """The 62-float records save_ply writes (io/ply.rs:67-98), little-endian."""
    g = np.asarray(g, f32).reshape(-1, 59)
    rec = np.zeros((len(g), 62), f32)
    rec[:, 0:3] = g[:, POS:POS + 3]
    rec[:, 6] = g[:, SH]
    rec[:, 7] = g[:, SH + 16]
    rec[:, 8] = g[:, SH + 32]
    k = 9
    for ch in range(3):
        for j in range(1, 16):
            rec[:, k] = g[:, SH + ch * 16 + j]
            k += 1
    rec[:, 54] = g[:, OPL]
    rec[:, 55:58] = g[:, LS:LS + 3]
    rec[:, 58:62] = g[:, ROT:ROT + 4]
    return rec.astype("<f4").tobytes()


def crafted_files(tmp):
    """(path, expected records) pairs for load_ply."""
    out = []

    # 1. ASCII, reordered properties, a missing value, CRLF, comments, an
    # unknown property and a face element whose list property is skipped
    p = tmp / "ascii.ply"
    p.write_bytes(b"ply\r\nformat ascii 1.0\r\ncomment made by hand\r\nelement vertex 3\r\n"
                  b"property float rot_3\r\nproperty float z\r\nproperty float red\r\nproperty float opacity\r\n"
                  b"property float f_rest_16\r\nelement face 0\r\nproperty list uchar int vertex_indices\r\n"
                  b"end_header\r\n"
                  b"0.25 -1.5e2 7 3\t 0.125\r\n"
                  b"1 2 3\r\n"
                  b"inf nan x1 +4.5 -0\r\n")
    e = np.stack([default_record() for _ in range(3)])
    e[0, ROT + 3], e[0, POS + 2], e[0, OPL], e[0, SH + 16 + 1 + 1] = 0.25, -150.0, 3.0, 0.125
    e[1, ROT + 3], e[1, POS + 2], e[1, OPL], e[1, SH + 17 + 1] = 1.0, 2.0, 0.0, 0.0  # missing -> 0
    e[2, ROT + 3], e[2, POS + 2], e[2, OPL], e[2, SH + 18] = np.inf, np.nan, 4.5, -0.0
    out.append((p, e))

    # 2. binary big-endian with double, uchar, ushort and int properties;
    # int is decoded as the bits of a float (io/ply.rs:196-199)
    p = tmp / "be.ply"
    hdr = (b"ply\nformat binary_big_endian 1.0\nelement vertex 2\nproperty double x\nproperty uchar scale_1\n"
           b"property ushort f_dc_2\nproperty int rot_0\nproperty float y\nend_header\n")
    recs = []
    vals = [(0.1, 200, 65535, 1.5, -2.0), (-1e300, 0, 7, -0.0, 3.25)]
    e = np.stack([default_record() for _ in range(2)])
    for i, (x, u8, u16, bits, y) in enumerate(vals):
        recs.append(np.array(x, ">f8").tobytes() + bytes([u8]) + np.array(u16, ">u2").tobytes()
                    + np.array(bits, ">f4").tobytes() + np.array(y, ">f4").tobytes())
        with np.errstate(over="ignore"):
            e[i, POS] = f32(x)
        e[i, LS + 1] = u8
        e[i, SH + 32] = u16
        e[i, ROT] = bits
        e[i, POS + 1] = y
    p.write_bytes(hdr + b"".join(recs))
    out.append((p, e))

    # 3. a degree-1 style f_rest list (9 entries) maps onto channel 0 at
    # 15 coefficients per channel (io/ply.rs:238-243), and a repeated
    # property: the later one wins
    p = tmp / "rest9.ply"
    names = [f"f_rest_{i}" for i in range(9)] + ["x", "x", "f_rest_44", "f_rest_45", "f_rest_+3", "f_rest_x"]
    hdr = ("ply\nformat binary_little_endian 1.0\nelement vertex 1\n"
           + "".join(f"property float {n}\n" for n in names) + "end_header\n").encode()
    v = np.arange(1, len(names) + 1, dtype=f32)
    p.write_bytes(hdr + v.astype("<f4").tobytes())
    e = default_record()[None].copy()
    e[0, SH + 1:SH + 10] = v[:9]
    e[0, POS] = v[10]
    e[0, SH + 32 + 15] = v[11]   # f_rest_44 -> channel 2, k 14
    e[0, SH + 4] = v[13]         # f_rest_+3 -> usize parse accepts '+'
    out.append((p, e))

    # 4. no vertex properties at all: defaults
    p = tmp / "empty_props.ply"
    p.write_bytes(b"ply\nformat binary_little_endian 1.0\nelement vertex 2\nend_header\n")
    out.append((p, np.stack([default_record(), default_record()])))

    # 5. zero vertices
    p = tmp / "zero.ply"
    p.write_bytes(b"ply\nformat binary_little_endian 1.0\nelement vertex 0\nproperty float x\nend_header\n")
    out.append((p, np.zeros((0, 59), f32)))
    return out


def bad_files(tmp):
    """Files load_ply rejects (io/ply.rs bails)."""
    files = {
        "magic": b"plx\nformat ascii 1.0\nelement vertex 0\nend_header\n",
        "format": b"ply\nformat binary_middle_endian 1.0\nelement vertex 0\nend_header\n",
        "type": b"ply\nformat ascii 1.0\nelement vertex 1\nproperty half x\nend_header\n1\n",
        "count": b"ply\nformat ascii 1.0\nelement vertex -1\nend_header\n",
        "eof": b"ply\nformat ascii 1.0\nelement vertex 1\n",
        "short": b"ply\nformat binary_little_endian 1.0\nelement vertex 2\nproperty float x\nend_header\n\0\0\0\0",
    }
    out = {}
    for k, b in files.items():
        p = tmp / f"bad_{k}.ply"
        p.write_bytes(b)
        out[k] = p
    return out
