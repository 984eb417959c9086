""".ptcl v0.2 files and the resume tooling around them.

Format (src/dtype.h:19-36, written by src/main.cpp:60-74,187-292): a packed
50-byte little-endian header "<4sBBIIIIIffffff" = ("PTCL", major 0,
minor 2, header_length 50, N, step, t_0, fps, x_min, x_max, y_min, y_max,
z_min, z_max) followed by one frame per time step: N x 3 float32
positions (or N x 12 with --force-data-output: position + 9 force floats).

Tools restated from the reference (same arguments, same results):
  generate_initialization_file  tools/generate_initialization_file.py:16-27
  split                         tools/split_datafile.py:11-31
  merge                         tools/merge_datafiles.py:11-22
  fix                           tools/fix_datafile.py:37-46
  calc_vel                      tools/swarm_util.py:195-220
Command line: python -m boidssimulatormpi_amd.ptcl {geninit,split,merge,fix,check} ...
"""
from __future__ import annotations

import os
import shutil
import struct
import sys

import numpy as np

HEADER_FORMAT = "<4sBBIIIIIffffff"
HEADER_SIZE = struct.calcsize(HEADER_FORMAT)   # 50


class Header:
    __slots__ = ("magic", "major", "minor", "header_length", "N", "steps", "t_0", "fps",
                 "x_min", "x_max", "y_min", "y_max", "z_min", "z_max")

    def __init__(self, N, steps, field, t_0=0, fps=30):
        self.magic, self.major, self.minor, self.header_length = b"PTCL", 0, 2, HEADER_SIZE
        self.N, self.steps, self.t_0, self.fps = N, steps, t_0, fps
        self.x_min = self.y_min = self.z_min = 0.0
        self.x_max = self.y_max = self.z_max = float(np.float32(field))

    def pack(self) -> bytes:
        return struct.pack(HEADER_FORMAT, *(getattr(self, k) for k in self.__slots__))

    @classmethod
    def unpack(cls, raw: bytes) -> "Header":
        h = cls.__new__(cls)
        for k, v in zip(cls.__slots__, struct.unpack(HEADER_FORMAT, raw[:HEADER_SIZE])):
            setattr(h, k, v)
        return h


class PtclFile:
    """Read-only view of a .ptcl file (dim = 3, or 12 with forces)."""

    def __init__(self, path, force=False):
        self.path = path
        self.dim = 12 if force else 3
        with open(path, "rb") as f:
            self.header = Header.unpack(f.read(HEADER_SIZE))
        self.N = self.header.N
        self.frame_bytes = 4 * self.dim * self.N
        self.file_size = os.path.getsize(path)
        self.valid_steps = (self.file_size - HEADER_SIZE) // self.frame_bytes if self.frame_bytes else 0
        self.correct_file_size = HEADER_SIZE + self.frame_bytes * self.valid_steps

    def check_metadata(self):
        """Raises ValueError when the header's frame count disagrees with the
        frames the file holds, or when a partial frame trails the data."""
        if self.valid_steps != self.header.steps:
            raise ValueError(f"frame count mismatch: the header declares {self.header.steps} frames, "
                             f"the file holds {self.valid_steps} whole frames")
        if (self.file_size - HEADER_SIZE) % self.frame_bytes:
            extra = self.file_size - self.correct_file_size
            raise ValueError(f"partial frame: {extra} bytes follow the last whole frame "
                             f"({self.file_size} bytes in the file, {self.correct_file_size} expected)")
        return True

    def frames(self, start=0, stop=None):
        stop = self.valid_steps if stop is None else stop
        if stop <= start:
            return np.zeros((0, self.N, self.dim), np.float32)
        with open(self.path, "rb") as f:
            f.seek(HEADER_SIZE + start * self.frame_bytes)
            raw = f.read(self.frame_bytes * (stop - start))
        return np.frombuffer(raw, "<f4").reshape(stop - start, self.N, self.dim).copy()

    def frame(self, t):
        return self.frames(t, t + 1)[0]


def write(path, frames, field, t_0=0, fps=30):
    """Write frames (T x N x 3 or T x N x 12 float32) as .ptcl v0.2."""
    frames = np.asarray(frames, np.float32)
    T, N = frames.shape[0], frames.shape[1]
    with open(path, "wb") as f:
        f.write(Header(N, T, field, t_0, fps).pack())
        f.write(frames.astype("<f4").tobytes())


def generate_initialization_file(pos_file, vel_file, time_step, out_file):
    """Interleave positions and velocities of frame `time_step` into the
    N x 6 float32 init file that init() reads (src/boid_simulation.cpp:113-136)."""
    N = PtclFile(pos_file).N
    with open(pos_file, "rb") as pf, open(vel_file, "rb") as vf:
        pf.seek(HEADER_SIZE + N * 12 * time_step)
        vf.seek(HEADER_SIZE + N * 12 * time_step)
        pos = pf.read(12 * N)
        vel = vf.read(12 * N)
    if len(pos) != 12 * N or len(vel) != 12 * N:
        raise ValueError("frame out of range")
    p = np.frombuffer(pos, "<f4").reshape(N, 3)
    v = np.frombuffer(vel, "<f4").reshape(N, 3)
    np.concatenate([p, v], 1).astype("<f4").tofile(out_file)


def split(in_file, out_file, start, steps):
    """Frames [start, start+steps) of in_file (absolute times, t_0 aware)."""
    src = PtclFile(in_file)
    t0_in = src.header.t_0
    h = src.header
    h.steps, h.t_0 = steps, start
    with open(in_file, "rb") as inf, open(out_file, "wb") as outf:
        outf.write(h.pack())
        inf.seek(HEADER_SIZE + src.frame_bytes * (start - t0_in))
        for k in range(steps):
            data = inf.read(src.frame_bytes)
            if len(data) != src.frame_bytes:
                raise ValueError(f"{in_file} ends before frame t={start + k} (holds t={t0_in}.."
                                 f"{t0_in + src.valid_steps - 1})")
            outf.write(data)


def fix(path):
    """Rewrite the header step count to the whole frames present and drop a
    trailing partial frame."""
    src = PtclFile(path)
    h = src.header
    h.steps = src.valid_steps
    with open(path, "r+b") as f:
        f.write(h.pack())
        f.truncate(src.correct_file_size)


def merge(in_files, out_file):
    """Concatenate the frames of in_files (headers of all but the first
    dropped), then fix the step count."""
    shutil.copyfile(in_files[0], out_file)
    with open(out_file, "ab") as outf:
        for fn in in_files[1:]:
            with open(fn, "rb") as inf:
                inf.seek(HEADER_SIZE)
                shutil.copyfileobj(inf, outf)
    fix(out_file)


def calc_vel(ptcl: PtclFile, time, step=1):
    """Position at `time` and the mean velocity over the next `step` frames,
    unwrapping periodic jumps larger than half the field."""
    size = ptcl.header.x_max - ptcl.header.x_min
    fr = ptcl.frames(time, time + step + 1)[:, :, :3].astype(np.float64)
    x0 = fr[0]
    offset = np.zeros_like(x0)
    for k in range(step):
        v = fr[k + 1] - fr[k]
        offset -= np.sign(v) * (np.abs(v) > size / 2) * size
    return x0, (fr[step] - x0 + offset) / step


def _main(argv):
    if len(argv) < 2:
        print(__doc__)
        return 2
    cmd, a = argv[1], argv[2:]
    if cmd == "geninit":
        generate_initialization_file(a[0], a[1], int(a[2]), a[3])
    elif cmd == "split":
        split(a[0], a[1], int(a[2]), int(a[3]))
    elif cmd == "merge":
        merge(a[:-1], a[-1])
    elif cmd == "fix":
        fix(a[0])
    elif cmd == "check":
        rc = 0
        for fn in a:
            p = PtclFile(fn)
            summary = f"{fn}: N={p.N}, first frame t={p.header.t_0}, {p.header.steps} frames in the header"
            try:
                p.check_metadata()
                print(f"{summary}; ok")
            except ValueError as e:
                print(f"{summary}; BAD ({e})")
                rc = 1
        return rc
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(_main(sys.argv))
