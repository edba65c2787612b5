"""Write y3g_photoz_table.inc: the photometric-redshift width table of the
Costanzi-2026 P-operator integrands, SIGMA_PHOTOZ_TABLE_t
(src/models/sigma_photoz_table_t.cuh:15-75): 120 values of sigma_z on a
uniform grid over z in [0.1, 0.9].

They are tabulated numbers the functor is defined by, so they are taken as
they are from the reference source (run where the reference tree is at
/root/reference); the generated file is committed.

    python y3_cluster_cpp_amd/csrc/gen_photoz_table.py
"""
import os
import re

SRC = "/root/reference/src/models/sigma_photoz_table_t.cuh"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "y3g_photoz_table.inc")


def main():
    with open(SRC) as f:
        src = f.read()
    i = src.index("SIGMA[120] =")
    j = src.index("{", i)
    k = src.index("}", j)
    v = re.findall(r"[-+]?\d+\.\d*(?:[eE][-+]?\d+)?", src[j:k])
    assert len(v) == 120, len(v)
    lines = ["// Generated by gen_photoz_table.py from SIGMA_PHOTOZ_TABLE_t",
             "// (sigma_photoz_table_t.cuh:15-75): sigma_z at 120 uniform z over",
             "// [0.1, 0.9].  Do not edit.",
             "Y3G_PHOTOZ_TABLE(kPhotozSigma, 120) = {"]
    for a in range(0, 120, 4):
        lines.append("  " + ", ".join(v[a:a + 4]) + ",")
    lines.append("};")
    with open(OUT, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
