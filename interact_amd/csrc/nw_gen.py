"""Generate nw_grad_gen.h: displacement gradients of the building blocks of
the Nikkhoo & Walter (2015) half-space triangular dislocation (GJI 201,
1117-1139), by symbolic differentiation of the published displacement
fields.  The reference evaluates the same method in
src/green/tdstresshs.f90 (strains) and src/green/tddisphs.f90
(displacements); here the strains are derived, not restated:

  adis_grad   gradient of the "incomplete" displacement of an angular
              dislocation in a full space (Comninou & Dundurs 1975; the
              reference's Angdisdisp, src/green/tddisphs.f90:468-551),
              in the angular dislocation's own frame
  fsc_grad    gradient of the harmonic free-surface correction of an
              angular dislocation whose apex is at depth a (the
              reference's Angdisdispfsc, src/green/tddisphs.f90:556-704)
  burg_grad   gradient of the triangle's Burgers function
              -Omega / (4 pi), Omega the solid angle subtended by the
              triangle (Van Oosterom & Strackee 1983 form, as in TDdispFS,
              src/green/tddisphs.f90:235-262), in the triangle's frame

Gradients are returned as g[i][j] = d u_i / d x_j.  The strain of the full
space triangular dislocation is the symmetric part of the sum of the three
adis_grad terms and burg_grad times the Burgers vector; the free-surface
part is the symmetric part of the fsc_grad differences.

Run from the repository root (needs sympy; the generated header is
committed, so building the library does not):

    python interact_amd/csrc/nw_gen.py
"""
import os

import sympy as sp
from sympy.printing.c import C99CodePrinter

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "nw_grad_gen.h")


class Printer(C99CodePrinter):
    """C printer: small integer powers as products, halves as sqrt."""

    def _print_Pow(self, expr):
        b, e = expr.base, expr.exp
        bs = self.parenthesize(b, 100)
        if e == 2:
            return f"({bs}*{bs})"
        if e == 3:
            return f"({bs}*{bs}*{bs})"
        if e == -1:
            return f"(1.0/{bs})"
        if e == -2:
            return f"(1.0/({bs}*{bs}))"
        if e == -3:
            return f"(1.0/({bs}*{bs}*{bs}))"
        if e == sp.Rational(1, 2):
            return f"sqrt({self._print(b)})"
        if e == -sp.Rational(1, 2):
            return f"(1.0/sqrt({self._print(b)}))"
        if e == sp.Rational(3, 2):
            return f"({bs}*sqrt({self._print(b)}))"
        if e == -sp.Rational(3, 2):
            return f"(1.0/({bs}*sqrt({self._print(b)})))"
        return super()._print_Pow(expr)


PR = Printer()


def emit(name, args, doc, exprs, out_names, out_decl="double g[3][3]"):
    """One C function computing the given expressions with CSE."""
    repl, red = sp.cse(exprs, symbols=sp.numbered_symbols("t"), optimizations="basic")
    lines = [f"/* {doc} */", f"NW_FN void {name}({', '.join('double ' + a for a in args)}, {out_decl})", "{"]
    for s, e in repl:
        lines.append(f"  const double {s} = {PR.doprint(e)};")
    for o, e in zip(out_names, red):
        lines.append(f"  {o} = {PR.doprint(e)};")
    lines.append("}")
    return "\n".join(lines) + "\n"


def grad(u, xs):
    return [sp.diff(ui, xj) for ui in u for xj in xs]


def adis(basis=False):
    """Incomplete full-space angular-dislocation displacement: apex at the
    origin, the dislocation line along -z bent by the angle whose cosine and
    sine are cA, sA; Burgers vector (bx, by, bz); Poisson ratio nu.  With
    basis=True the gradients for the three unit Burgers vectors."""
    x, y, z, cA, sA, bx, by, bz, nu = sp.symbols("x y z cA sA bx by bz nu", real=True)
    eta = y * cA - z * sA
    zeta = y * sA + z * cA
    r = sp.sqrt(x**2 + y**2 + z**2)
    q = 1 / (8 * sp.pi * (1 - nu))
    m = 1 - 2 * nu
    Lz, Lzeta = sp.log(r - z), sp.log(r - zeta)
    rz, rzeta = r * (r - z), r * (r - zeta)
    u = bx * q * (x * y / rz - x * eta / rzeta) \
        + by * q * (x**2 * cA / rzeta - x**2 / rz - m * (cA * Lzeta - Lz)) \
        + bz * sA * q * (m * Lzeta - x**2 / rzeta)
    v = bx * q * (eta * sA / (r - zeta) - y * eta / rzeta + y**2 / rz + m * (cA * Lzeta - Lz)) \
        + by * x * q * (y * cA / rzeta - sA * cA / (r - zeta) - y / rz) \
        + bz * x * sA * q * (sA / (r - zeta) - y / rzeta)
    w = bx * q * (eta * cA / (r - zeta) - y / r - eta * z / rzeta - m * sA * Lzeta) \
        + by * x * q * (z * cA / rzeta - cA**2 / (r - zeta) + 1 / r) \
        + bz * x * sA * q * (cA / (r - zeta) - z / rzeta)
    if basis:
        out = []
        for c in range(3):
            sub = {bx: int(c == 0), by: int(c == 1), bz: int(c == 2)}
            out += grad([e.subs(sub) for e in (u, v, w)], [x, y, z])
        return ["x", "y", "z", "cA", "sA", "nu"], out
    return ["x", "y", "z", "cA", "sA", "bx", "by", "bz", "nu"], grad([u, v, w], [x, y, z])


def fsc(basis=False):
    """Free-surface correction of an angular dislocation with apex depth a
    (image apex at y3 = -a), angle with cosine cB and sine sB.  With
    basis=True the gradients for the three unit Burgers vectors."""
    y1, y2, y3, cB, sB, b1, b2, b3, nu, a = sp.symbols("y1 y2 y3 cB sB b1 b2 b3 nu a", real=True)
    ct = cB / sB                   # cot(beta)
    cth = -sB / (cB - 1)           # cot(beta / 2)
    Y3 = y3 + 2 * a
    Z1 = y1 * cB + Y3 * sB
    Z3 = -y1 * sB + Y3 * cB
    R2 = y1**2 + y2**2 + Y3**2
    R = sp.sqrt(R2)
    F = 2 * sp.atan2(-y2, -(R + Y3) * cth + y1)     # image Burgers function
    m, n2, c2 = 1 - 2 * nu, 2 * nu, 2 * (1 - nu)
    ar = a / R
    RY, RZ = R + Y3, R + Z3
    LY, LZ = sp.log(R + Y3), sp.log(R + Z3)
    Ym = Y3 - a
    K = 1 / (4 * sp.pi * (1 - nu))
    rcb = R * cB + Y3
    # Burgers component b1
    v1b1 = (-c2 * m * F * ct**2 + m * y2 / RY * ((m - ar) * ct - y1 / RY * (nu + ar))
            + m * y2 * cB * ct / RZ * (cB + ar) + a * y2 * Ym * ct / R**3
            + y2 * Ym / (R * RY) * (-m * ct + y1 / RY * (n2 + ar) + a * y1 / R2)
            + y2 * Ym / (R * RZ) * (cB / RZ * (rcb * (m * cB - ar) * ct + c2 * (R * sB - y1) * cB)
                                    - a * Y3 * cB * ct / R2))
    v2b1 = (m * ((c2 * ct**2 - nu) * LY - (c2 * ct**2 + m) * cB * LZ)
            - m / RY * (y1 * ct * (m - ar) + nu * Y3 - a + y2**2 / RY * (nu + ar))
            - m * Z1 * ct / RZ * (cB + ar) - a * y1 * Ym * ct / R**3
            + Ym / RY * (-n2 + (m * y1 * ct - a) / R + y2**2 / (R * RY) * (n2 + ar) + a * y2**2 / R**3)
            + Ym / RZ * (cB**2 - (m * Z1 * ct + a * cB) / R + a * Y3 * Z1 * ct / R**3
                         - (y2**2 * cB**2 - a * Z1 * ct / R * rcb) / (R * RZ)))
    v3b1 = (c2 * (m * F * ct + y2 / RY * (n2 + ar) - y2 * cB / RZ * (cB + ar))
            + y2 * Ym / R * (n2 / RY + a / R2)
            + y2 * Ym * cB / (R * RZ) * (m - rcb / RZ * (cB + ar) - a * Y3 / R2))
    # Burgers component b2
    v1b2 = (m * ((c2 * ct**2 + nu) * LY - (c2 * ct**2 + 1) * cB * LZ)
            + m / RY * (-m * y1 * ct + nu * Y3 - a + a * y1 * ct / R + y1**2 / RY * (nu + ar))
            - m * ct / RZ * (Z1 * cB - a * (R * sB - y1) / (R * cB)) - a * y1 * Ym * ct / R**3
            + Ym / RY * (n2 + (m * y1 * ct + a) / R - y1**2 / (R * RY) * (n2 + ar) - a * y1**2 / R**3)
            + Ym * ct / RZ * (-cB * sB + a * y1 * Y3 / (R**3 * cB)
                              + (R * sB - y1) / R * (c2 * cB - rcb / RZ * (1 + a / (R * cB)))))
    v2b2 = (c2 * m * F * ct**2 + m * y2 / RY * (-(m - ar) * ct + y1 / RY * (nu + ar))
            - m * y2 * ct / RZ * (1 + a / (R * cB)) - a * y2 * Ym * ct / R**3
            + y2 * Ym / (R * RY) * (m * ct - n2 * y1 / RY - a * y1 / R * (1 / R + 1 / RY))
            + y2 * Ym * ct / (R * RZ) * (-c2 * cB + rcb / RZ * (1 + a / (R * cB)) + a * Y3 / (R2 * cB)))
    v3b2 = (-c2 * m * ct * (LY - cB * LZ) - c2 * y1 / RY * (n2 + ar) + c2 * Z1 / RZ * (cB + ar)
            + Ym / R * (m * ct - n2 * y1 / RY - a * y1 / R2)
            - Ym / RZ * (cB * sB + rcb * ct / R * (c2 * cB - rcb / RZ)
                         + ar * (sB - Y3 * Z1 / R2 - Z1 * rcb / (R * RZ))))
    # Burgers component b3
    v1b3 = (m * (y2 / RY * (1 + ar) - y2 * cB / RZ * (cB + ar)) - y2 * Ym / R * (a / R2 + 1 / RY)
            + y2 * Ym * cB / (R * RZ) * (rcb / RZ * (cB + ar) + a * Y3 / R2))
    v2b3 = (m * (-sB * LZ - y1 / RY * (1 + ar) + Z1 / RZ * (cB + ar)) + y1 * Ym / R * (a / R2 + 1 / RY)
            - Ym / RZ * (sB * (cB - ar) + Z1 / R * (1 + a * Y3 / R2)
                         - (y2**2 * cB * sB - a * Z1 / R * rcb) / (R * RZ)))
    v3b3 = (c2 * F + c2 * (y2 * sB / RZ * (cB + ar))
            + y2 * Ym * sB / (R * RZ) * (1 + rcb / RZ * (cB + ar) + a * Y3 / R2))
    v1 = K * (b1 * v1b1 + b2 * v1b2 + b3 * v1b3)
    v2 = K * (b1 * v2b1 + b2 * v2b2 + b3 * v2b3)
    v3 = K * (b1 * v3b1 + b2 * v3b2 + b3 * v3b3)
    if basis:
        out = []
        for c in range(3):
            sub = {b1: int(c == 0), b2: int(c == 1), b3: int(c == 2)}
            out += grad([e.subs(sub) for e in (v1, v2, v3)], [y1, y2, y3])
        return ["y1", "y2", "y3", "cB", "sB", "nu", "a"], out
    return ["y1", "y2", "y3", "cB", "sB", "b1", "b2", "b3", "nu", "a"], grad([v1, v2, v3], [y1, y2, y3])


def burgers():
    """-Omega / (4 pi) of the triangle (0, p1y, p1z), (0, 0, 0), (0, p3y, p3z)
    seen from (x, y, z), the triangle's frame with the second vertex at the
    origin; returned as a 1 x 3 gradient in g[0]."""
    x, y, z, p1y, p1z, p3y, p3z = sp.symbols("x y z p1y p1z p3y p3z", real=True)
    A = sp.Matrix([-x, p1y - y, p1z - z])
    B = sp.Matrix([-x, -y, -z])
    C = sp.Matrix([-x, p3y - y, p3z - z])
    na, nb, nc = (sp.sqrt(V.dot(V)) for V in (A, B, C))
    num = A.dot(B.cross(C))
    den = na * nb * nc + A.dot(B) * nc + A.dot(C) * nb + B.dot(C) * na
    fi = -2 * sp.atan2(num, den) / (4 * sp.pi)
    return ["x", "y", "z", "p1y", "p1z", "p3y", "p3z"], [sp.diff(fi, s) for s in (x, y, z)]


def main():
    parts = [
        "/* Generated by nw_gen.py from the published displacement fields; do not edit.\n"
        "   Displacement gradients g[i][j] = d u_i / d x_j of the angular-dislocation\n"
        "   building blocks of the Nikkhoo & Walter (2015) triangular dislocation.\n"
        "   Compiles as C (the CPU oracle) and as HIP device code (NW_FN set to\n"
        "   __device__ __forceinline__ before inclusion). */\n",
        "#pragma once\n#include <math.h>\n#ifndef NW_FN\n#define NW_FN static inline\n#endif\n",
    ]
    a, e = adis()
    parts.append(emit("nw_adis_grad", a,
                      "full-space angular dislocation, incomplete displacement (Comninou & Dundurs 1975)",
                      e, [f"g[{i}][{j}]" for i in range(3) for j in range(3)]))
    a, e = fsc()
    parts.append(emit("nw_fsc_grad", a,
                      "free-surface correction of an angular dislocation, apex depth a",
                      e, [f"g[{i}][{j}]" for i in range(3) for j in range(3)]))
    b3 = [f"g[{c}][{i}][{j}]" for c in range(3) for i in range(3) for j in range(3)]
    a, e = adis(basis=True)
    parts.append(emit("nw_adis_gradb", a, "nw_adis_grad for the unit Burgers vectors: g[c] for b = e_c",
                      e, b3, "double g[3][3][3]"))
    a, e = fsc(basis=True)
    parts.append(emit("nw_fsc_gradb", a, "nw_fsc_grad for the unit Burgers vectors: g[c] for b = e_c",
                      e, b3, "double g[3][3][3]"))
    a, e = burgers()
    parts.append(emit("nw_burg_grad", a, "Burgers function -Omega/(4 pi) of the triangle, gradient in g[0]",
                      e, [f"g[0][{j}]" for j in range(3)]))
    with open(OUT, "w") as f:
        f.write("\n".join(parts))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
