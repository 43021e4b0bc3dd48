"""Write tests/golden/sh_reference_basis.npz: the reference's SH basis and its partials.

Reads msplat/src/compute_sh.cu of the reference as text (never compiles or
imports it), turns every forward term ``sh[k] * SH_Cl[i] * <poly>`` into an
exact sympy polynomial, and evaluates Y_k and dY_k/d(x,y,z) in float64 at
seeded directions.  Those are the values the reference's forward computes and
the partials its backward (compute_sh.cu:505-1598) returns as dL/dsh and
dL/ddir per unit upstream gradient.  Directions include unit vectors and a
few non-unit ones (the partials of the literal polynomials differ there).

Usage: python tests/golden/make_sh_golden.py [/path/to/reference]
"""

import os
import re
import sys

import numpy as np
import sympy as sp

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sh_reference_basis.npz")


def parse(path):
    src = open(path).read()
    x, y, z = sp.symbols("x y z", real=True)
    consts = {"SH_C0": [sp.Float("0.28209479177387814", 30)]}
    for name, body in re.findall(r"__device__ const float (SH_C\d+)\[\] = \{([^}]*)\}", src):
        consts[name] = [sp.Float(v.strip().rstrip("f"), 30) for v in body.split(",") if v.strip()]
    loc = {"x": x, "y": y, "z": z, "xy": x * y, "yz": y * z, "xz": x * z, "xyz": x * y * z}
    for a, s in (("x", x), ("y", y), ("z", z)):
        for p in (2, 3, 4, 6, 8, 10):
            loc[f"{a}{p}"] = s ** p
    fwd = src[: src.index("Backward(")]
    polys = {0: consts["SH_C0"][0]}
    for m in re.finditer(r"result\s*\+=\s*(.*?);", fwd, re.S):
        e = m.group(1)
        k = int(re.search(r"sh\[(\d+)\]", e).group(1))
        e = re.sub(r"sh\[\d+\]", "1", e)
        e = re.sub(r"(\d+\.\d*)f", r"\1", e)  # C float literal suffix
        e = re.sub(r"(SH_C\d+)\[(\d+)\]", lambda mm: "(" + str(consts[mm.group(1)][int(mm.group(2))]) + ")", e)
        polys[k] = sp.expand(sp.sympify(e, locals=loc))
    return (x, y, z), polys


def main():
    (x, y, z), polys = parse(os.path.join(REF, "msplat", "src", "compute_sh.cu"))
    assert sorted(polys) == list(range(121)), sorted(polys)
    rng = np.random.default_rng(20240611)
    d = rng.normal(size=(24, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[20:] *= np.array([0.7, 1.3, 0.9, 1.1])[:, None]  # four non-unit directions
    Y = np.zeros((len(d), 121))
    dY = np.zeros((len(d), 121, 3))
    for k in range(121):
        p = polys[k]
        fs = [sp.lambdify((x, y, z), e, "math") for e in (p, sp.diff(p, x), sp.diff(p, y), sp.diff(p, z))]
        for i, (a, b, c) in enumerate(d):
            Y[i, k] = float(fs[0](a, b, c))
            dY[i, k] = [float(f(a, b, c)) for f in fs[1:]]
    np.savez_compressed(OUT, dirs=d, Y=Y, dY=dY)
    print("wrote", OUT, Y.shape, dY.shape)


if __name__ == "__main__":
    main()
