"""Packs the CBCL face images that GPUMLib's FacesNMF demo trains on
(src/examples/FacesNMF/images/face00001.pgm ... face02429.pgm, 19 x 19 P5) into
tests/golden/cbcl_faces.npz: the exact bytes of every file (data, uint8, all files
concatenated in order) and where each file starts (offsets, int64, one more than
the file count).  Data only: the tests write the files back out and load them with
the drop-in driver, so the GPU box, which has no copy of the reference tree, sees
the same inputs.  The files are kept byte for byte as the reference tree holds them:
2400 have LF line ends, 29 have CR LF ones (two of those also carry a CR LF in
their raster), and the loader's reading of each kind is part of what is tested.
Run from the repository root, with the reference tree present:
    python tests/golden/make_cbcl_faces.py [path/to/FacesNMF/images]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT = "/root/reference/src/examples/FacesNMF/images"
COUNT = 2429


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else DEFAULT
    blobs = []
    for k in range(1, COUNT + 1):
        with open(os.path.join(src, f"face{k:05d}.pgm"), "rb") as f:
            blobs.append(f.read())
    offsets = np.zeros(COUNT + 1, np.int64)
    offsets[1:] = np.cumsum([len(b) for b in blobs])
    data = np.frombuffer(b"".join(blobs), np.uint8)
    out = os.path.join(HERE, "cbcl_faces.npz")
    np.savez_compressed(out, data=data, offsets=offsets)
    print(f"{out}: {COUNT} files, {data.size} bytes")


if __name__ == "__main__":
    main()
