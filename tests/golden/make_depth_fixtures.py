"""Writes tests/golden/imrod_v{1,2}.depth.gz: the two .depth files the reference's
own tests read (KinectFission/content/, used by UnitTests/Integrator.cpp:30 and
UnitTests/Mesher.cpp:32), gzip-compressed as they are.  imrod_v1.depth is a
version-1 file (640 x 480 uint16 mm) whose header promises 18 frames but which
holds one; imrod_v2.depth is a version-2 file with one frame of binary32 metres.
Run in a checkout that has the reference at /root/reference."""
import gzip
import os
import shutil

SRC = "/root/reference/KinectFission/content"
HERE = os.path.dirname(os.path.abspath(__file__))

if __name__ == "__main__":
    for name in ("imrod_v1.depth", "imrod_v2.depth"):
        with open(os.path.join(SRC, name), "rb") as f, \
                gzip.GzipFile(os.path.join(HERE, name + ".gz"), "wb", compresslevel=9, mtime=0) as g:
            shutil.copyfileobj(f, g)
