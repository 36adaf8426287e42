"""Write tests/golden/slots.json: the first-triplet slot sizes of index_t
(size/offset/max_size per direction, initial_sort, sufsort.c:386-450) that the
reference's own construct_index computes for the golden reads files, both index
widths (oracle/_ref/ref_dump_*, which prints them).  Run from the repository
root after `make -C oracle`:

    python tests/golden/make_slot_golden.py
"""
import gzip
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pq_oracle as orc  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    cases = json.load(open(os.path.join(HERE, "cases.json")))
    tmp = tempfile.mkdtemp()
    out = {}
    for name, meta in sorted(cases.items()):
        fa = os.path.join(tmp, name + ".fa")
        with gzip.open(os.path.join(HERE, name + ".fa.gz"), "rb") as fi, open(fa, "wb") as fo:
            shutil.copyfileobj(fi, fo)
        out[name] = {}
        for width in ("u32", "u64"):
            exe = os.path.join(orc.REFDIR, f"ref_dump_{width}")
            r = subprocess.run([exe, fa, str(meta["L"]), str(meta["t"]),
                                os.path.join(tmp, "d.bin"), "nograph"],
                               capture_output=True, text=True, check=True,
                               env=dict(os.environ, OMP_NUM_THREADS="1"))
            m = re.search(r"slots fw ([0-9 ]+) bw ([0-9 ]+)", r.stdout)
            v = [int(x) for x in m.group(1).split()] + [int(x) for x in m.group(2).split()]
            out[name][width] = {"size_fw": v[0:5], "offset_fw": v[5:10], "max_size_fw": v[10],
                                "size_bw": v[11:16], "offset_bw": v[16:21], "max_size_bw": v[21]}
            print(name, width, out[name][width])
    with open(os.path.join(HERE, "slots.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    shutil.rmtree(tmp)


if __name__ == "__main__":
    main()
