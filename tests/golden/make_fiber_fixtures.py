"""Write the lattice and b-table fixtures the tests use, from the data files
the reference ships (hpc_code/input/*/fiber.mat, hpc_code/btable/*.mat).

Run in a checkout that has the reference next to it:
    python tests/golden/make_fiber_fixtures.py /root/reference
Each fixture is one axon of a reference dataset, x/y cropped to the non-null
bounding box plus 2 voxels (full z), saved compressed with its voxel size.
Only data is kept; nothing of the reference's code is copied.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

DATASETS = [
    ("I_realistic_axon_mitochondria", 0),
    ("II_realistic_axon", 0),
    ("III_caliber_variation", 0),
    ("IV_undulation", 0),
]


def main(ref):
    import __graft_entry__ as ge
    rms = ge.load_package()
    import importlib
    fileio = importlib.import_module(rms.__name__ + ".fileio")
    for name, k in DATASETS:
        fib = fileio.load_fiber_mat(os.path.join(ref, "hpc_code", "input", name, "fiber.mat"))
        lab, vs = fib[k]
        crop = fileio.crop_to_content(lab, margin=2)
        out = os.path.join(HERE, f"fiber_{name.split('_')[0]}_{k:03d}.npz")
        np.savez_compressed(out, labels=crop.astype(np.uint8), voxel_size=np.float64(vs),
                            full_shape=np.array(lab.shape), count=np.int64(len(fib)))
        print(out, crop.shape, np.bincount(crop.ravel()).tolist(), vs)
    from scipy.io import loadmat
    bt = loadmat(os.path.join(ref, "hpc_code", "btable", "btable_5shell_90ndir_b1000.mat"))
    out = os.path.join(HERE, "btable_5shell_90ndir_b1000.npz")
    np.savez_compressed(out, bval=np.asarray(bt["bval"], float).ravel(), bvec=np.asarray(bt["bvec"], float))
    print(out, np.unique(np.round(np.asarray(bt["bval"]).ravel(), 6)))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
