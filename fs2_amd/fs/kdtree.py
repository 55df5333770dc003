"""Copy of the kd-tree of the last FoF run (py/fs/kdtree.py of the reference)."""
from . import _fs as c


class KdTree:
    """kdtree[i] = (i, k, left, right, ibegin, iend): node i holds particles
    [ibegin, iend) whose coordinate k spans [left, right]."""

    def __init__(self):
        self._list = c._kdtree_create_copy()
        self.height = c._kdtree_get_height()

    def __getitem__(self, idx):
        return self._list[idx]

    def __len__(self):
        return len(self._list)

    @staticmethod
    def left_child(i):
        return 2 * i + 1

    @staticmethod
    def right_child(i):
        return 2 * i + 2
