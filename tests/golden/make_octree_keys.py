"""Extract the reference's octree-key test fixtures as data.

The reference's CPU cuckoo test (leven/src/test_cuckoo.cpp:120-179) inserts the
uint32 octree node codes of leven/src/testdata/octree_keys_*.cpp (each list
terminated by ~0).  This script reads those arrays and stores the key values in
tests/golden/octree_keys.npz (one uint32 array per fixture, terminator
dropped).  Run once in a container where the reference tree is mounted:

    python tests/golden/make_octree_keys.py /root/reference/leven/src/testdata
"""
import glob
import os
import re
import sys

import numpy as np


def main(src_dir):
    out = {}
    for path in sorted(glob.glob(os.path.join(src_dir, "octree_keys_*.cpp"))):
        name = os.path.splitext(os.path.basename(path))[0]
        text = open(path).read()
        vals = [int(v, 16) for v in re.findall(r"0x[0-9a-fA-F]+", text)]
        keys = np.array(vals, dtype=np.uint64)
        keys = keys[keys != 0xFFFFFFFF].astype(np.uint32)
        out[name] = keys
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "octree_keys.npz")
    np.savez_compressed(dst, **out)
    print(dst, {k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/leven/src/testdata")
