"""Write y3g_lc_lt_tables.inc: the tables of the Costanzi et al. 2018
eq. 15 richness transfer LC_LT_t (src/models/lc_lt_t.cc:10-214): the knots
lt_bins and zt_bins without their last (upper-limit) entries, as the
reference's make_short_vec drops them (22 lambda_true and 5 z_true knots),
and the five 22 x 5 tables tau, mu, sigma, fmsk, fprj, lambda_true fastest.

They are fitted numbers the functor is defined by, so they are taken as they
are from the reference source (run where the reference tree is at
/root/reference); the generated file is committed.

    python y3_cluster_cpp_amd/csrc/gen_lc_lt_tables.py
"""
import os
import re

SRC = "/root/reference/src/models/lc_lt_t.cc"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "y3g_lc_lt_tables.inc")


def braced_numbers(src, name):
    """The numbers of the brace initialiser that follows `name =`."""
    i = src.index(name + " =")
    j = src.index("{", i)
    k = src.index("}", j)
    body = re.sub(r"//[^\n]*", "", src[j:k])
    return [x for x in re.findall(r"[-+]?\d+\.\d*(?:[eE][-+]?\d+)?", body)]


def main():
    with open(SRC) as f:
        src = f.read()
    arrays = [("kLcLtLt", "lt_bins", 22), ("kLcLtZt", "zt_bins", 5),
              ("kLcLtTau", "tau_arr", 110), ("kLcLtMu", "mu_arr", 110),
              ("kLcLtSigma", "sigma_arr", 110), ("kLcLtFmsk", "fmsk_arr", 110),
              ("kLcLtFprj", "fprj_arr", 110)]
    lines = ["// Generated by gen_lc_lt_tables.py from the tables of LC_LT_t",
             "// (lc_lt_t.cc:10-214): 22 lambda_true knots, 5 z_true knots (the",
             "// reference's bins without their upper limits) and the tables tau, mu,",
             "// sigma, fmsk, fprj [zt knot][lt knot].  Do not edit."]
    for cname, rname, n in arrays:
        v = braced_numbers(src, rname)
        assert len(v) == n, (rname, len(v))
        lines.append(f"Y3G_LC_LT_TABLE({cname}, {n}) = {{")
        for k in range(0, n, 4):
            lines.append("  " + ", ".join(v[k:k + 4]) + ",")
        lines.append("};")
    with open(OUT, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
