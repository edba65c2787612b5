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
"""
import os

SRC = "/root/reference/data"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_data")

PLAN = [("test_Rmis.txt", 2, 1), ("test_dv_do_dz.txt", 0, 1),
        ("z_da_test.txt", 0, 1), ("d_a_test.txt", 0, 1),
        ("test_int_zo_zt_des.txt", 2, 1), ("test_zo_zt_des.txt", 2, 4),
        ("test_lolc_t.txt", 2, 50)]


def main():
    os.makedirs(DST, exist_ok=True)
    for name, nheader, stride in PLAN:
        with open(os.path.join(SRC, name)) as f:
            lines = f.read().splitlines()
        head, body = lines[:nheader], [l for l in lines[nheader:] if l.strip()]
        keep = body[::stride]
        with open(os.path.join(DST, name), "w") as f:
            f.write("\n".join(head + keep) + "\n")
        print(f"{name}: {len(keep)} of {len(body)} rows")


if __name__ == "__main__":
    main()
