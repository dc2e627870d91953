"""Common base for leaf profiles: how a leaf enters the lowered k-space program."""
import numpy as np

from . import _lib
from .gsobject import GSObject


def jac_record(jac):
    """(separable, [dudx, dudy, dvdx, dvdy]) for SBProfile::drawK's three cases
    (src/SBProfile.cpp:399-424): no Jacobian, diagonal, general."""
    if jac is None:
        return 1, (1., 0., 0., 1.)
    j = np.asarray(jac, dtype=float).reshape(2, 2)
    if j[0, 1] == 0. and j[1, 0] == 0.:
        return 1, (j[0, 0], 0., 0., j[1, 1])
    return 0, (j[0, 0], j[0, 1], j[1, 0], j[1, 1])


class Leaf(GSObject):
    _leaf_type = 0

    def _fill(self, rec, ctx, jac):
        raise NotImplementedError

    def _lower(self, ctx, jac):
        from .render import LeafRecord
        rec = LeafRecord()
        sep, j = jac_record(jac)
        rec['type'] = self._leaf_type
        rec['separable'] = sep
        rec['jac'] = j
        rec['cliprows'] = -1
        self._fill(rec, ctx, jac)
        ctx.emit(_lib.OP_LEAF, ctx.add_leaf(rec))
