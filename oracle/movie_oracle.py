"""CPU restatement of the reference's movie writer (test infrastructure: only
tests/ import it, as the checker of the library's movie frames).

Follows DSMC_IO::save_state_to_movie_file (dsmc_io.cpp:50-80), which
main.cpp:57-59 calls before every step: when create_movie is set and the
step count is a multiple of movie_every_n_frame, the frame buffer (6 floats
per molecule for MAX_MOLECULE_NUM molecules, zeroed once when the file is
opened, dsmc_io.cpp:53-60) takes r and v of the molecules in storage order
as float32 (62-69), and the file gets the int movie_molecules / ranks
(here one rank) and that many rows of the buffer (72-75).  Pinned in
tests/test_movie.py against the compiled reference's own writer
(oracle/_ref) on the same states, including a count above the number of
molecules (rows of the zeroed buffer).
"""
import numpy as np


class MovieWriter:
    def __init__(self, count, capacity):
        """count: movie_molecules (one rank); capacity: MAX_MOLECULE_NUM
        (max_molecules_per_node, system.cpp:244)."""
        if not 0 <= count <= capacity:
            raise ValueError("the reference reads past its frame buffer for count > capacity")
        self.count = int(count)
        self.buf = np.zeros((capacity, 6), np.float32)
        self.data = bytearray()

    def before_step(self, steps, every, rv6):
        """main.cpp:58: the writer before the step whose count is `steps`."""
        if steps % every:
            return
        rv6 = np.asarray(rv6, np.float64)
        self.buf[:len(rv6)] = rv6.astype(np.float32)
        self.data += np.int32(self.count).tobytes() + self.buf[:self.count].tobytes()


def read_frames(data):
    """The frames of a movie file's bytes: a list of (count, 6) float32."""
    out, pos = [], 0
    while pos < len(data):
        n = int(np.frombuffer(data, np.int32, 1, pos)[0])
        pos += 4
        out.append(np.frombuffer(data, np.float32, 6 * n, pos).reshape(n, 6).copy())
        pos += 24 * n
    return out
