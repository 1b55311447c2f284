"""Reader and writer of .depth files (reference: KinectFission/kifi/DepthStream.cpp:15-102;
format: poly2depth/README.txt).

Layout, little endian: a 15-byte magic, int32 version; version 2 adds int32
width, height and texel type (0 = uint16 mm, 1 = binary32 metres); version 1 is
640 x 480 uint16.  Then int32 frame count, and per frame 16 float32 of view
matrix followed by width x height texels, row-major.  The view is returned as
the 16 floats taken as column storage: the reference reads them through
float4x4's row-major constructor and transposes once more for the file's
row-vector convention (DepthStream.cpp:69-71).
"""
import numpy as np

MAGIC = b"KPPL raw depth\n"
SHORT, FLOAT = 0, 1


class DepthStream:
    """DepthStream (DepthStream.h:11-45).  next_frame() returns (depth, view) or None
    once FrameCount frames are read, or where the file ends early (the reference
    does not check its stream there)."""

    def __init__(self, path):
        self._f = open(path, "rb")
        head = self._f.read(15)
        if len(head) != 15:
            raise ValueError("not a .depth file: %s" % path)
        version = self._int()
        if version == 1:
            self.width, self.height, self.texel = 640, 480, SHORT
        elif version == 2:
            self.width, self.height, self.texel = self._int(), self._int(), self._int()
            if self.texel not in (SHORT, FLOAT) or self.width <= 0 or self.height <= 0:
                raise ValueError("bad .depth header")
        else:
            raise ValueError("unsupported .depth version %d" % version)
        self.version = version
        self.frames = self._int()
        self.index = 0

    def _int(self):
        b = self._f.read(4)
        if len(b) != 4:
            raise ValueError("truncated .depth header")
        return int(np.frombuffer(b, "<i4")[0])

    def frame_count(self):
        return self.frames

    def next_frame(self, metres=True):
        """(depth, view float32[16] column storage).  metres: float32 metres as the
        reference returns them (uint16 texels as mm * 0.001f, DepthStream.cpp:84-91);
        otherwise the texels as stored (uint16 mm or float32 m)."""
        if self.index >= self.frames:
            return None
        vb = self._f.read(64)
        n = self.width * self.height
        tb = self._f.read(n * (2 if self.texel == SHORT else 4))
        if len(vb) != 64 or len(tb) != n * (2 if self.texel == SHORT else 4):
            return None
        view = np.frombuffer(vb, "<f4").astype(np.float32)
        raw = np.frombuffer(tb, "<u2" if self.texel == SHORT else "<f4").reshape(self.height, self.width)
        self.index += 1
        if metres and self.texel == SHORT:
            return (raw.astype(np.float32) * np.float32(0.001)).astype(np.float32), view
        return raw.astype(np.uint16 if self.texel == SHORT else np.float32), view

    def close(self):
        self._f.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def write_depth(path, frames, views, version=2):
    """Writes frames ((H, W) uint16 mm or float32 metres, all alike) and their views
    (float32[16] column storage) as a .depth file; version 1 needs 640 x 480 uint16."""
    frames = [np.asarray(f) for f in frames]
    H, W = frames[0].shape
    texel = FLOAT if frames[0].dtype == np.float32 else SHORT
    if version == 1 and (texel != SHORT or (W, H) != (640, 480)):
        raise ValueError("version 1 holds 640 x 480 uint16 frames only")
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(np.array([version], "<i4").tobytes())
        if version == 2:
            f.write(np.array([W, H, texel], "<i4").tobytes())
        f.write(np.array([len(frames)], "<i4").tobytes())
        for d, v in zip(frames, views):
            if d.shape != (H, W):
                raise ValueError("frames differ in size")
            f.write(np.asarray(v, "<f4").reshape(16).tobytes())
            f.write(np.ascontiguousarray(d, "<u2" if texel == SHORT else "<f4").tobytes())
