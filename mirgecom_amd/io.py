"""File-name helpers of mirgecom/io.py:75-82."""


def make_rank_fname(basename, rank=0, step=0, t=0):
    """Per-rank visualization file name pattern (``{rank}`` left for
    ``str.format``)."""
    return f"{basename}-{step:09d}-{{rank:04d}}.vtu"


def make_par_fname(basename, step=0, t=0):
    """Parallel visualization manifest name."""
    return f"{basename}-{step:09d}.pvtu"
