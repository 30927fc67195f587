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


def oracle_to_product_tag(tag):
    from oracle import tags as OT
    from carcassonne_amd import sparse as PT
    if isinstance(tag, OT.Identity):
        return PT.Identity()
    if isinstance(tag, OT.Complete):
        return PT.Complete()
    if isinstance(tag, OT.OneSiteOperator):
        return PT.OneSiteOperator(None)
    if isinstance(tag, OT.TwoSiteOperator):
        return PT.TwoSiteOperator(tag.id, tag.direction, tag.position)
    if isinstance(tag, OT.TwoSiteOperatorCompressed):
        return PT.TwoSiteOperatorCompressed(tag.direction)
    raise KeyError(tag)


def product_to_oracle_tag(tag):
    from oracle import tags as OT
    from carcassonne_amd import sparse as PT
    if isinstance(tag, PT.Identity):
        return OT.Identity()
    if isinstance(tag, PT.Complete):
        return OT.Complete()
    if isinstance(tag, PT.OneSiteOperator):
        return OT.OneSiteOperator(None)
    if isinstance(tag, PT.TwoSiteOperator):
        return OT.TwoSiteOperator(tag.id, tag.direction, tag.position)
    if isinstance(tag, PT.TwoSiteOperatorCompressed):
        return OT.TwoSiteOperatorCompressed(tag.direction)
    raise KeyError(tag)


def system_to_product(osys):
    """The same tensors as an oracle System, as a product System on the GPU."""
    from carcassonne_amd.data import HipData
    from carcassonne_amd.system import System
    up = HipData.fromArray
    return System([{oracle_to_product_tag(t): up(v) for t, v in c.items()} for c in osys.corners],
                  [{oracle_to_product_tag(t): up(v) for t, v in s.items()} for s in osys.sides],
                  up(osys.state), {oracle_to_product_tag(t): up(v) for t, v in osys.op_center.items()})


def compare_systems(psys, osys, rtol=1e-9):
    """Largest relative difference over every corner / side channel and the state."""
    worst = 0.0
    for pl, ol in ((psys.corners, osys.corners), (psys.sides, osys.sides)):
        for pd, od in zip(pl, ol):
            assert set(product_to_oracle_tag(t) for t in pd) == set(od.keys()), (list(pd), list(od))
            for t, v in pd.items():
                w = od[product_to_oracle_tag(t)]
                assert v.shape == w.shape, (t, v.shape, w.shape)
                worst = max(worst, rel_err(v.toArray(), w))
    worst = max(worst, rel_err(psys.state_center_data.toArray(), osys.state))
    return worst
