"""Test helpers: convert synthetic environments / product tensors to the oracle's tags and arrays."""
import numpy as np

from oracle import tags as OT


def oracle_tag(key):
    k = key[0]
    if k == "I":
        return OT.Identity()
    if k == "C":
        return OT.Complete()
    if k == "T":
        return OT.TwoSiteOperator(key[1], OT.CENTER)
    if k == "O":
        return OT.OneSiteOperator(None)
    if k == "TT":
        return OT.TwoSiteOperator(key[1], key[2], 0)
    raise KeyError(key)


def product_tag(key):
    from carcassonne_amd import sparse as PT
    k = key[0]
    if k == "I":
        return PT.Identity()
    if k == "C":
        return PT.Complete()
    if k == "T":
        return PT.TwoSiteOperator(key[1], PT.CENTER)
    if k == "O":
        return PT.OneSiteOperator(None)
    if k == "TT":
        return PT.TwoSiteOperator(key[1], key[2], 0)
    raise KeyError(key)


def to_oracle(corners, sides, op):
    return ([{oracle_tag(k): v for k, v in c.items()} for c in corners],
            [{oracle_tag(k): v for k, v in s.items()} for s in sides],
            {oracle_tag(k): v for k, v in op.items()})


def to_product(corners, sides, op):
    from carcassonne_amd.data import HipData
    up = HipData.fromArray
    return ([{product_tag(k): up(v) for k, v in c.items()} for c in corners],
            [{product_tag(k): up(v) for k, v in s.items()} for s in sides],
            {product_tag(k): up(v) for k, v in op.items()})


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
