"""Colour initialisation of scene files: every scene colour becomes Xyz<E>.

Restates clovers/src/colorinit.rs:45-60 and the background conversion of
clovers-cli/src/scenefile.rs:43-44, which use the `palette` 0.7.6 crate:

* legacy arrays ``[r, g, b]`` and ``{"srgb": ...}`` are *encoded* sRGB, decoded
  with the sRGB EOTF (IEC 61966-2-1), also above 1.0 (an emitter of 25 is
  ~2002 linear);
* linear sRGB -> XYZ(D65) uses the matrix derived from the sRGB primaries
  (x,y) = R(0.64,0.33) G(0.30,0.60) B(0.15,0.06) and the D65 white point
  (0.95047, 1.0, 1.08883);
* XYZ(D65) -> XYZ(E) uses Bradford chromatic adaptation (Lam 1985) with the
  von Kries scaling in the Bradford cone space.

palette evaluates these in f32; this module evaluates them in f64 and rounds
once to f32.  No reference test pins palette's rounding, so the last-ulp
agreement of these conversions is not pinned (closed forms are tested).
"""
import numpy as np

SRGB_PRIMARIES_XY = ((0.64, 0.33), (0.30, 0.60), (0.15, 0.06))
WHITE_D65 = np.array([0.95047, 1.0, 1.08883])
WHITE_E = np.array([1.0, 1.0, 1.0])
BRADFORD = np.array([[0.8951, 0.2664, -0.1614],
                     [-0.7502, 1.7135, 0.0367],
                     [0.0389, -0.0685, 1.0296]])


def srgb_eotf(c):
    """sRGB transfer function, decoding (IEC 61966-2-1)."""
    c = np.asarray(c, dtype=np.float64)
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def rgb_to_xyz_matrix(primaries=SRGB_PRIMARIES_XY, white=WHITE_D65):
    """RGB->XYZ matrix from primary chromaticities and the white point XYZ."""
    cols = []
    for x, y in primaries:
        cols.append([x / y, 1.0, (1.0 - x - y) / y])
    P = np.array(cols).T
    S = np.linalg.solve(P, white)
    return P * S


def bradford_matrix(src_white=WHITE_D65, dst_white=WHITE_E):
    """XYZ(src) -> XYZ(dst) via Bradford cone response scaling."""
    rho_s = BRADFORD @ src_white
    rho_d = BRADFORD @ dst_white
    return np.linalg.inv(BRADFORD) @ np.diag(rho_d / rho_s) @ BRADFORD


M_SRGB_TO_XYZ_D65 = rgb_to_xyz_matrix()
M_D65_TO_E = bradford_matrix()


def linsrgb_to_xyz_e(rgb):
    return M_D65_TO_E @ (M_SRGB_TO_XYZ_D65 @ np.asarray(rgb, dtype=np.float64))


def srgb_to_xyz_e(rgb):
    return linsrgb_to_xyz_e(srgb_eotf(rgb))


def xyz_d65_to_xyz_e(xyz):
    return M_D65_TO_E @ np.asarray(xyz, dtype=np.float64)


def _triple(v, names=("red", "green", "blue")):
    if isinstance(v, dict):
        return [float(v[n]) for n in names]
    return [float(a) for a in v]


def color_init_to_xyz_e(c):
    """ColorInit (untagged: legacy array or typed) -> Xyz<E> as float32[3]."""
    if isinstance(c, (list, tuple)):
        out = srgb_to_xyz_e(_triple(c))
    elif isinstance(c, dict) and len(c) == 1:
        (tag, val), = c.items()
        if tag == "hex":
            h = val.lstrip("#")
            if len(h) == 3:
                h = "".join(ch * 2 for ch in h)
            rgb = [int(h[i:i + 2], 16) / 255.0 for i in (0, 2, 4)]
            out = srgb_to_xyz_e(rgb)
        elif tag == "lin_srgb":
            out = linsrgb_to_xyz_e(_triple(val))
        elif tag == "srgb":
            out = srgb_to_xyz_e(_triple(val))
        elif tag == "xyz_e":
            out = np.array(_triple(val, ("x", "y", "z")))
        elif tag == "xyz_d65":
            out = xyz_d65_to_xyz_e(_triple(val, ("x", "y", "z")))
        else:
            raise NotImplementedError("colour space %r is not supported" % tag)
    else:
        raise ValueError("unrecognised colour %r" % (c,))
    return np.asarray(out, dtype=np.float32)
