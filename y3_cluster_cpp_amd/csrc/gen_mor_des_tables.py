"""Write y3g_mor_des_tables.inc: the skew-normal scatter tables of the DES
mass-observable relation MOR_DES_t (src/models/mor_des_t.cuh:11-214), the
knots test_sigintr (14) and test_lsat (42) and the 14 x 42 tables
sig_skewnorml_flat and skews, row = sigma_intr knot, column = lambda_sat knot.

They are fitted numbers the functor is defined by, so they are taken as they
are from the reference header (run where the reference tree is at
/root/reference); the generated file is committed.

    python y3_cluster_cpp_amd/csrc/gen_mor_des_tables.py
"""
import os
import re

SRC = "/root/reference/src/models/mor_des_t.cuh"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "y3g_mor_des_tables.inc")


def braced_numbers(src, name):
    """The numbers of the brace initialiser that follows `name`."""
    i = src.index(name)
    j = src.index("{", i)
    depth, k = 0, j
    while True:
        if src[k] == "{":
            depth += 1
        elif src[k] == "}":
            depth -= 1
            if depth == 0:
                break
        k += 1
    body = re.sub(r"//[^\n]*", "", src[j:k])
    return [float(x) for x in re.findall(r"[-+]?\d+\.\d*(?:[eE][-+]?\d+)?", body)]


def main():
    with open(SRC) as f:
        src = f.read()
    arrays = [("kMorDesSigintr", "test_sigintr", 14), ("kMorDesLsat", "test_lsat", 42),
              ("kMorDesSigma", "sig_skewnorml_flat", 14 * 42),
              ("kMorDesSkew", "skews", 14 * 42)]
    lines = ["// Generated by gen_mor_des_tables.py from the skew-normal scatter tables",
             "// of MOR_DES_t (mor_des_t.cuh:11-214): knots in sigma_intr (14) and",
             "// lambda_sat (42); tables [sigma_intr knot][lambda_sat knot].",
             "// Do not edit."]
    for cname, rname, n in arrays:
        v = braced_numbers(src, rname)
        assert len(v) == n, (rname, len(v))
        lines.append(f"Y3G_MOR_DES_TABLE({cname}, {n}) = {{")
        for k in range(0, n, 6):
            lines.append("  " + ", ".join(f"{x:.6f}" for x in v[k:k + 6]) + ",")
        lines.append("};")
    with open(OUT, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
