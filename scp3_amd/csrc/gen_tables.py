"""Regenerate the numeric tables the kernels need from the reference sources.

The tables are data, not code: the Bratley-Fox / Joe-Kuo primitive
polynomials and initial direction numbers used by SCP3's Sobol generator
(sobol.cpp:14308-14421 for ``poly``, sobol.cpp:14441-27676 for the initial
``v[i][j]``), and the Partridge-Schwenke monomer potential / dipole surface
coefficients used by TTM3-F (ps.cpp).  They must match the reference
exactly for the Sobol integers to be bit-identical, so they are extracted
rather than retyped.

Usage:  python scp3_amd/csrc/gen_tables.py [/path/to/reference]

The generated ``.inc`` files are committed; this script is only rerun when
the tables need to be refreshed.
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _sobol(ref):
    src = open(os.path.join(ref, "sobol.cpp")).read()
    # the i8_sobol copy of the polynomial table (the i4 copy is identical
    # but we take the one used by the 64-bit generator)
    start = src.index("void i8_sobol")
    body = src[start:]
    m = re.search(r"static long long int poly\[DIM_MAX2\] =\s*\{([^}]*)\}", body)
    poly = [int(t) for t in m.group(1).replace("\n", " ").split(",") if t.strip()]
    assert len(poly) == 1111, len(poly)
    init = {}
    end = body.index("Check parameters.")
    for i, j, val in re.findall(r"v\[(\d+)\]\[(\d+)\] = (\d+);", body[:end]):
        init[(int(i), int(j))] = int(val)
    degs = []
    for d in range(1111):
        p = poly[d]
        m_ = 0
        while p > 1:
            p >>= 1
            m_ += 1
        degs.append(m_)
    maxdeg = max(degs)
    rows = []
    for d in range(1111):
        rows.append([init.get((d, j), 0) for j in range(maxdeg)])
    out = []
    out.append("// Generated by gen_tables.py from the reference sobol.cpp "
               "(poly: sobol.cpp:14308-14421, initial v: sobol.cpp:14441-27676).")
    out.append("// Bratley-Fox / Joe-Kuo primitive polynomials and initial direction numbers.")
    out.append("#define SCP3_SOBOL_DIM_MAX 1111")
    out.append("#define SCP3_SOBOL_INIT_MAX %d" % maxdeg)
    out.append("static const unsigned short scp3_sobol_poly[SCP3_SOBOL_DIM_MAX] = {")
    for k in range(0, 1111, 16):
        out.append("  " + ", ".join(str(x) for x in poly[k:k + 16]) + ",")
    out.append("};")
    out.append("static const unsigned short scp3_sobol_init[SCP3_SOBOL_DIM_MAX][SCP3_SOBOL_INIT_MAX] = {")
    for r in rows:
        out.append("  {" + ", ".join(str(x) for x in r) + "},")
    out.append("};")
    return "\n".join(out) + "\n"


def _ps(ref):
    """Coefficient and exponent tables of ps.cpp as data.

    c5z is combined as the reference does at first use (ps.cpp:507-512),
    ((f5z*c5zA + fbasis*cbasis) + fcore*ccore) + frest*crest, in IEEE double
    (Python floats), so the values are bit-identical.
    """
    src = open(os.path.join(ref, "ps.cpp")).read()
    arrays = {}
    pat = re.compile(r"const\s+(double|size_t)\s+(\w+)\[(\d+)\]\s*=\s*\{(.*?)\};", re.S)
    for typ, name, n, body in pat.findall(src):
        vals = [t.strip() for t in body.replace("\n", " ").split(",") if t.strip()]
        assert len(vals) == int(n), (name, len(vals), n)
        arrays[name] = (typ, vals)
    sc = {}
    for nm in ("f5z", "fbasis", "fcore", "frest"):
        sc[nm] = float(re.search(r"const double %s = ([-0-9.eE+]+);" % nm, src).group(1))
    A = [float(v) for v in arrays["c5zA"][1]]
    Bb = [float(v) for v in arrays["cbasis"][1]]
    Cc = [float(v) for v in arrays["ccore"][1]]
    Rr = [float(v) for v in arrays["crest"][1]]
    c5z = [sc["f5z"] * A[i] + sc["fbasis"] * Bb[i] + sc["fcore"] * Cc[i] + sc["frest"] * Rr[i]
           for i in range(245)]
    out = ["// Generated by gen_tables.py from the reference ps.cpp: Partridge-Schwenke",
           "// monomer potential (c5z = f5z*c5zA + fbasis*cbasis + fcore*ccore + frest*crest,",
           "// ps.cpp:507-512, exponent tables idx1..3) and dipole surface (coefD, idxD0..2).",
           "// SCP3_PS_STORAGE lets C++/HIP code make them constexpr.",
           "#ifndef SCP3_PS_STORAGE",
           "#define SCP3_PS_STORAGE static const",
           "#endif"]
    out.append("SCP3_PS_STORAGE double scp3_ps_c5z[245] = {")
    for k in range(0, 245, 4):
        out.append("  " + ", ".join(repr(v) for v in c5z[k:k + 4]) + ",")
    out.append("};")
    for name in ("idx1", "idx2", "idx3", "idxD0", "idxD1", "idxD2"):
        vals = arrays[name][1]
        out.append("SCP3_PS_STORAGE unsigned char scp3_ps_%s[%d] = {" % (name, len(vals)))
        for k in range(0, len(vals), 20):
            out.append("  " + ", ".join(vals[k:k + 20]) + ",")
        out.append("};")
    vals = arrays["coefD"][1]
    out.append("SCP3_PS_STORAGE double scp3_ps_coefD[84] = {")
    for k in range(0, 84, 3):
        out.append("  " + ", ".join(vals[k:k + 3]) + ",")
    out.append("};")
    return "\n".join(out) + "\n"


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(HERE, "sobol_dirnums.inc"), "w") as f:
        f.write(_sobol(ref))
    if os.path.exists(os.path.join(ref, "ps.cpp")):
        with open(os.path.join(HERE, "ps_coeffs.inc"), "w") as f:
            f.write(_ps(ref))


if __name__ == "__main__":
    main()
