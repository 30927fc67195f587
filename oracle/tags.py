"""Operator-channel tags (oracle).  Restates carcassonne/sparse.py:9-345.

A sparse tensor is a dict tag -> dense ndarray.  Tags carry the bookkeeping of
which operator terms a tensor already contains.  Semantics follow the
reference exactly, including its quirk that OneSiteOperator ignores its id
(sparse.py:36-37), so all one-site tags compare equal.

Pinned by: the closed-form walks (identity operator E = width*height, Z field
product states) and the 2D-in-1D known answers in tests/test_oracle_*.py,
which fail if any tag rule drops or duplicates a term.
"""

LEFT, RIGHT, CENTER = 0, 1, 2          # sparse.py:211-215


def O(i): return (i + 2) % 4             # utils.py:886
def L(i): return (i + 1) % 4             # utils.py:887
def R(i): return (i - 1) % 4             # utils.py:888


class _Singleton:
    _instance = None

    def __new__(cls):
        if cls.__dict__.get("_instance") is None:
            cls._instance = object.__new__(cls)
        return cls._instance

    def __hash__(self):
        return hash(type(self).__name__)

    def __repr__(self):
        return type(self).__name__ + "()"


class Identity(_Singleton):
    _instance = None


class Complete(_Singleton):
    _instance = None


class OneSiteOperator:
    """sparse.py:34-54 -- the id is discarded on construction."""
    __slots__ = ["id"]

    def __init__(self, id=None):
        self.id = None

    def __eq__(self, other):
        return isinstance(other, OneSiteOperator) and other.id == self.id

    def __hash__(self):
        return hash(("OneSiteOperator", self.id))

    def __repr__(self):
        return "OneSiteOperator(None)"


class TwoSiteOperator:
    """sparse.py:55-170."""
    __slots__ = ["id", "direction", "position"]

    def __init__(self, id, direction, position=None):
        self.id, self.direction, self.position = id, direction, position

    def __eq__(self, other):
        return (isinstance(other, TwoSiteOperator) and other.id == self.id
                and other.direction == self.direction and other.position == self.position)

    def __hash__(self):
        return hash(("TwoSiteOperator", self.id, self.direction, self.position))

    def __repr__(self):
        return "TwoSiteOperator({},{},{})".format(self.id, self.direction, self.position)

    def with_(self, direction, position=None):
        return TwoSiteOperator(self.id, direction, position)

    def moveOut(self):
        return self.with_(self.direction, self.position + 1)

    # each rule returns the result tag or None (sparse.py:81-166)
    def matches(left, right):
        if (left.id == right.id and left.direction == RIGHT and right.direction == LEFT
                and left.position == right.position):
            return Complete()

    def matchesCenter(self, side_direction, center):
        if self.id == center.id and self.direction == CENTER and side_direction == center.direction:
            return Complete()

    def matchesCenterIdentity(self):
        if self.direction != CENTER:
            return self.moveOut()

    def matchesCenterForStage3(self, direction, other):
        return (self.id == other.id and self.direction == CENTER
                and self.position + 2 * direction == other.direction)

    def matchesCornerIdentityOnLeft(self):
        if self.direction == RIGHT:
            return self
        if self.direction == CENTER:
            return self.with_(LEFT, 0)

    def matchesCornerIdentityOnLeftForStage1(self):
        if self.direction in (RIGHT, CENTER):
            return self

    def matchesCornerIdentityOnRight(self):
        if self.direction == LEFT:
            return self
        if self.direction == CENTER:
            return self.with_(RIGHT, 0)

    def matchesSideIdentityOnLeft(self):
        if self.direction == RIGHT:
            return self.moveOut()

    def matchesSideIdentityOnRight(self):
        if self.direction == LEFT:
            return self.moveOut()

    def matchesSideIdentityOnRightForStage1(self):
        if self.direction == LEFT:
            return self

    def matchesSideIdentityOutward(self, direction):
        if self.direction == L(direction):
            return self.with_(LEFT, 0)
        if self.direction == R(direction):
            return self.with_(RIGHT, 0)
        if self.direction == O(direction):
            return self.with_(CENTER)

    def matchesStage1IdentityOnLeft(self):
        if self.direction == RIGHT:
            return self
        if self.direction == CENTER:
            return self.with_(CENTER, LEFT)

    def matchesStage1IdentityOnRight(self):
        if self.direction == LEFT:
            return self
        if self.direction == CENTER:
            return self.with_(CENTER, RIGHT)

    def matchesForStage3(x, y):
        return (x.id == y.id and (x.direction, y.direction) in ((LEFT, RIGHT), (RIGHT, LEFT))
                and x.position == y.position)


class TwoSiteOperatorCompressed:
    """sparse.py:171-208."""
    __slots__ = ["direction"]

    def __init__(self, direction):
        self.direction = direction

    def __eq__(self, other):
        return isinstance(other, TwoSiteOperatorCompressed) and other.direction == self.direction

    def __hash__(self):
        return hash(("TwoSiteOperatorCompressed", self.direction))

    def __repr__(self):
        return "TwoSiteOperatorCompressed({})".format(self.direction)

    def matches(left, right):
        if left.direction == RIGHT and right.direction == LEFT:
            return Complete()

    def matchesCornerIdentityOnLeftForStage1(self):
        if self.direction == RIGHT:
            return self

    def matchesSideIdentityOnRightForStage1(self):
        if self.direction == LEFT:
            return self

    def matchesStage1IdentityOnLeft(self):
        if self.direction == RIGHT:
            return self

    def matchesStage1IdentityOnRight(self):
        if self.direction == LEFT:
            return self

    def matchesForStage3(x, y):
        return (x.direction, y.direction) in ((LEFT, RIGHT), (RIGHT, LEFT))


def standard_terms(fn):
    """addStandardCompleteAndIdentityTerms (sparse.py:218-222)."""
    return {
        (Identity, Identity): lambda x, y: (Identity(), fn),
        (Complete, Identity): lambda x, y: (Complete(), fn),
        (Identity, Complete): lambda x, y: (Complete(), fn),
    }


def contract_sparse(rules, t1, t2):
    """contractSparseTensors (sparse.py:223-239): every tag pair whose type pair has a
    rule yields (result tag, dense fn); equal result tags are summed."""
    out = {}
    for tag1, d1 in t1.items():
        for tag2, d2 in t2.items():
            key = (type(tag1), type(tag2))
            if key not in rules:
                continue
            handler = rules[key](tag1, tag2)
            if handler is None:
                continue
            rtag, fn = handler
            if rtag is None:
                continue
            res = fn(d1, d2)
            if rtag in out:
                out[rtag] = out[rtag] + res
            else:
                out[rtag] = res
    return out


def make_sparse_operator(Os=(), OO_UDs=(), OO_LRs=()):
    """makeSparseOperator (sparse.py:295-338)."""
    import numpy as np
    op = {}
    ident = None
    shape = None

    def check(s):
        if shape != s:
            raise ValueError("incompatible site matrix shapes: {} and {}".format(shape, s))
    for i, o in enumerate(Os):
        op[OneSiteOperator(i)] = o
        if ident is None:
            ident, shape = np.identity(o.shape[0], dtype=o.dtype), o.shape
        else:
            check(o.shape)
    for i, (ou, od) in enumerate(OO_UDs):
        op[TwoSiteOperator(i, 3, 0)] = ou
        op[TwoSiteOperator(i, 1, 0)] = od
        if ident is None:
            ident, shape = np.identity(ou.shape[0], dtype=ou.dtype), ou.shape
        else:
            check(ou.shape)
        check(od.shape)
    for i, (ol, orr) in enumerate(OO_LRs):
        op[TwoSiteOperator(i, 0, 0)] = ol
        op[TwoSiteOperator(i, 2, 0)] = orr
        if ident is None:
            ident, shape = np.identity(ol.shape[0], dtype=ol.dtype), ol.shape
        else:
            check(ol.shape)
        check(orr.shape)
    if ident is None:
        raise ValueError("No terms have been specified.")
    op[Identity()] = ident
    return op


def make_simple_sparse_operator(O=None, OO_UD=None, OO_LR=None):
    """makeSimpleSparseOperator (sparse.py:288-293)."""
    return make_sparse_operator([O] if O is not None else [],
                                [OO_UD] if OO_UD is not None else [],
                                [OO_LR] if OO_LR is not None else [])


def strip_all_but_identity(sparse):
    return {Identity(): sparse[Identity()]}
