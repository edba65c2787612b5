"""Copy the reference's own known-answer data files (used by its Catch2 tests)
into tests/golden/ref_data, subsampling the large ones deterministically.

Run in a container that has the reference tree at /root/reference.  The
outputs are data (inputs and expected outputs), committed next to this
script; nothing else of the reference is copied.

    file                        reference test that reads it          kept
    data/test_Rmis.txt          test/roffset_t.test.cc:12-54          all
    data/test_dv_do_dz.txt      test/dv_do_dz_t.test.cc:17-52         all
    data/z_da_test.txt          test/dv_do_dz_t.test.cc:35            all
    data/d_a_test.txt           test/dv_do_dz_t.test.cc:37            all
    data/test_int_zo_zt_des.txt test/int_zo_zt_des_t.test.cc:15-59    all
    data/test_zo_zt_des.txt     test/zo_zt_des_t.test.cc:15-54        every 4th row
    data/test_lolc_t.txt        test/lo_lc_t.test.cc:13-58            every 50th row
    data/int_lc_lt_des_t_*.txt  test/int_lc_lt_des_t_test.test.cu     all (the
                                (tables of INT_LC_LT_DES_t2)          lt / zt knots
                                                                      and 4 tables)
    data/test_lc_lt_t_SDSS.txt  test/lc_lt_t.test.cc:11-51            all
    data/nfw_off_center/table_1000_1e-03_5e+03_gamma_logx.txt,
    ..._gamma_logxmis.txt       the axes NFW_DSIGMA_MIS reads          all
                                (src/models/nfw_dsigma_mis.cuh:29-41,
                                test/nfw_dsigma_mis.test.cu)

and the expected values of test/int_lc_lt_des_t_test.test.cu (its `lc`,
`lt`, `zt` lists and its 720-entry `answers` array, in the test's order
lc-major, then zt, then lt) as int_lc_lt_des_answers.json.
"""
import json
import os
import re

SRC = "/root/reference/data"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_data")

PLAN = [("test_Rmis.txt", 2, 1), ("test_dv_do_dz.txt", 0, 1),
        ("z_da_test.txt", 0, 1), ("d_a_test.txt", 0, 1),
        ("test_int_zo_zt_des.txt", 2, 1), ("test_zo_zt_des.txt", 2, 4),
        ("test_lolc_t.txt", 2, 50),
        ("int_lc_lt_des_t_lt_bins.txt", 0, 1), ("int_lc_lt_des_t_zt_bins.txt", 0, 1),
        ("int_lc_lt_des_t_lbd_bin0.txt", 0, 1), ("int_lc_lt_des_t_lbd_bin1.txt", 0, 1),
        ("int_lc_lt_des_t_lbd_bin2.txt", 0, 1), ("int_lc_lt_des_t_lbd_bin3.txt", 0, 1),
        ("test_lc_lt_t_SDSS.txt", 0, 1),
        ("nfw_off_center/table_1000_1e-03_5e+03_gamma_logx.txt", 0, 1),
        ("nfw_off_center/table_1000_1e-03_5e+03_gamma_logxmis.txt", 0, 1)]
LCLT_TEST = "/root/reference/test/int_lc_lt_des_t_test.test.cu"


def _numbers_of(src, name):
    """The numbers of the brace list assigned to std::vector<double> `name`."""
    m = re.search(r"std::vector<double>\s+" + name + r"\s*\{(.*?)\};", src, re.S)
    body = re.sub(r"//[^\n]*", "", m.group(1))
    return [float(x) for x in re.findall(r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?", body)]


def lclt_answers():
    with open(LCLT_TEST) as f:
        src = f.read()
    out = {k: _numbers_of(src, k) for k in ("lc", "lt", "zt", "answers")}
    assert len(out["answers"]) == len(out["lc"]) * len(out["lt"]) * len(out["zt"])
    with open(os.path.join(DST, "int_lc_lt_des_answers.json"), "w") as f:
        json.dump(out, f)
    print(f"int_lc_lt_des_answers.json: {len(out['answers'])} values")


def main():
    os.makedirs(DST, exist_ok=True)
    for name, nheader, stride in PLAN:
        with open(os.path.join(SRC, name)) as f:
            lines = f.read().splitlines()
        head, body = lines[:nheader], [l for l in lines[nheader:] if l.strip()]
        keep = body[::stride]
        with open(os.path.join(DST, os.path.basename(name)), "w") as f:
            f.write("\n".join(head + keep) + "\n")
        print(f"{name}: {len(keep)} of {len(body)} rows")
    lclt_answers()


if __name__ == "__main__":
    main()
