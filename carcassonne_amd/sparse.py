"""Operator-channel tags and sparse (tag -> data) tensors.  Re-implements
carcassonne/sparse.py:9-345 (tags, the sparse contraction loop and the operator
builders) with the same names, equality rules and result tags.

A sparse tensor is a dict tag -> Data.  These dicts are host-side bookkeeping;
the Data values live in device memory.
"""
LEFT, RIGHT, CENTER = 0, 1, 2


def _L(i): return (i + 1) % 4
def _R(i): return (i - 1) % 4
def _O(i): return (i + 2) % 4


class _Unique:
    _instance = None

    def __new__(cls):
        inst = cls.__dict__.get("_instance")
        if inst is None:
            inst = object.__new__(cls)
            cls._instance = inst
        return inst

    def __hash__(self):
        return hash(type(self).__name__)

    def __repr__(self):
        return type(self).__name__ + "()"


class Identity(_Unique):
    _instance = None


class Complete(_Unique):
    _instance = None


class OneSiteOperator:
    """All one-site tags compare equal: the id is dropped on construction (sparse.py:36-37)."""
    __slots__ = ["id"]

    def __init__(self, id=None):
        self.id = None

    def __eq__(self, other):
        return isinstance(other, OneSiteOperator) and other.id == self.id

    def __hash__(self):
        return hash((OneSiteOperator.__name__, self.id))

    def __repr__(self):
        return "OneSiteOperator({})".format(self.id)


class TwoSiteOperator:
    __slots__ = ["id", "direction", "position"]

    def __init__(self, id, direction, position=None):
        self.id, self.direction, self.position = id, direction, position

    def __eq__(self, other):
        return (isinstance(other, TwoSiteOperator) and (other.id, other.direction, other.position)
                == (self.id, self.direction, self.position))

    def __hash__(self):
        return hash((TwoSiteOperator.__name__, self.id, self.direction, self.position))

    def __repr__(self):
        return "TwoSiteOperator({},{},{})".format(self.id, self.direction, self.position)

    def withNewDirectionAndPosition(self, direction, position=None):
        return TwoSiteOperator(self.id, direction, position)

    def moveOut(self):
        return self.withNewDirectionAndPosition(self.direction, self.position + 1)

    def matches(left, right):
        if (left.id, left.direction, left.position) == (right.id, RIGHT, right.position) and right.direction == LEFT:
            return Complete()

    def matchesCenter(self, side_direction, center):
        if self.id == center.id and self.direction == CENTER and side_direction == center.direction:
            return Complete()

    def matchesCenterIdentity(self):
        if self.direction != CENTER:
            return self.moveOut()

    def matchesCenterForStage3(self, direction, other):
        return self.id == other.id and self.direction == CENTER and self.position + 2 * direction == other.direction

    def matchesCornerIdentityOnLeft(self):
        if self.direction == RIGHT:
            return self
        if self.direction == CENTER:
            return self.withNewDirectionAndPosition(LEFT, 0)

    def matchesCornerIdentityOnLeftForStage1(self):
        if self.direction in (RIGHT, CENTER):
            return self

    def matchesCornerIdentityOnRight(self):
        if self.direction == LEFT:
            return self
        if self.direction == CENTER:
            return self.withNewDirectionAndPosition(RIGHT, 0)

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
        if self.direction == _L(direction):
            return self.withNewDirectionAndPosition(LEFT, 0)
        if self.direction == _R(direction):
            return self.withNewDirectionAndPosition(RIGHT, 0)
        if self.direction == _O(direction):
            return self.withNewDirectionAndPosition(CENTER)

    def matchesStage1IdentityOnLeft(self):
        if self.direction == RIGHT:
            return self
        if self.direction == CENTER:
            return self.withNewDirectionAndPosition(CENTER, LEFT)

    def matchesStage1IdentityOnRight(self):
        if self.direction == LEFT:
            return self
        if self.direction == CENTER:
            return self.withNewDirectionAndPosition(CENTER, RIGHT)

    def matchesForStage3(x, y):
        return x.id == y.id and {x.direction, y.direction} == {LEFT, RIGHT} and x.position == y.position


class TwoSiteOperatorCompressed:
    __slots__ = ["direction"]

    def __init__(self, direction):
        self.direction = direction

    def __eq__(self, other):
        return isinstance(other, TwoSiteOperatorCompressed) and other.direction == self.direction

    def __hash__(self):
        return hash((TwoSiteOperatorCompressed.__name__, self.direction))

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
        return {x.direction, y.direction} == {LEFT, RIGHT}


def addStandardCompleteAndIdentityTerms(terms, dense):
    terms[Identity, Identity] = lambda x, y: (Identity(), dense)
    terms[Complete, Identity] = lambda x, y: (Complete(), dense)
    terms[Identity, Complete] = lambda x, y: (Complete(), dense)


def contractSparseTensors(dense_contractors, tensor_1, tensor_2):
    """Pairs every tag of tensor_1 with every tag of tensor_2 through the rule table
    (type1, type2) -> handler(tag1, tag2) -> (result tag, dense fn); results with
    equal tags are accumulated in place."""
    result = {}
    for tag_1, data_1 in tensor_1.items():
        for tag_2, data_2 in tensor_2.items():
            rule = dense_contractors.get((type(tag_1), type(tag_2)))
            if rule is None:
                continue
            handler = rule(tag_1, tag_2)
            if handler is None or handler[0] is None:
                continue
            result_tag, contract = handler
            data = contract(data_1, data_2)
            if result_tag in result:
                result[result_tag] += data
            else:
                result[result_tag] = data
    return result


def formSparseContractor(dense_contractors):
    return lambda t1, t2: contractSparseTensors(dense_contractors, t1, t2)


def getInformationFromOperatorCenter(operator_center):
    for matrix in operator_center.values():
        return matrix.shape[0], matrix.dtype, type(matrix)
    raise ValueError("operator tensor has no term from which to extract the physical dimension")


def makeSparseOperator(Os=(), OO_UDs=(), OO_LRs=()):
    """sparse.py:295-338: one-site terms, up/down pairs at directions (3, 1), left/right
    pairs at (0, 2), plus the identity."""
    operator, identity, shape = {}, None, None

    def register(o):
        nonlocal identity, shape
        if identity is None:
            identity, shape = o.newIdentity(o.shape[0]), o.shape
        elif o.shape != shape:
            raise ValueError("incompatible site matrix shapes: {} and {}".format(shape, o.shape))
    for i, o in enumerate(Os):
        operator[OneSiteOperator(i)] = o
        register(o)
    for i, (o_u, o_d) in enumerate(OO_UDs):
        operator[TwoSiteOperator(i, 3, 0)] = o_u
        operator[TwoSiteOperator(i, 1, 0)] = o_d
        register(o_u)
        register(o_d)
    for i, (o_l, o_r) in enumerate(OO_LRs):
        operator[TwoSiteOperator(i, 0, 0)] = o_l
        operator[TwoSiteOperator(i, 2, 0)] = o_r
        register(o_l)
        register(o_r)
    if identity is None:
        raise ValueError("No terms have been specified.")
    operator[Identity()] = identity
    return operator


def makeSimpleSparseOperator(O=None, OO_UD=None, OO_LR=None):
    return makeSparseOperator([O] if O is not None else [], [OO_UD] if OO_UD is not None else [],
                              [OO_LR] if OO_LR is not None else [])


def mapOverSparseData(f, sparse):
    return {tag: f(data) for tag, data in sparse.items()}


def stripAllButIdentityFrom(sparse):
    return {Identity(): sparse[Identity()]}


def directSumSparse(sparse1, sparse2):
    z = tuple(s[Identity()].newZeros(s[Identity()].shape) for s in (sparse1, sparse2))
    out = {}
    for tag in list(sparse1.keys()) + [t for t in sparse2.keys() if t not in sparse1]:
        a = sparse1.get(tag, z[0])
        b = sparse2.get(tag, z[1])
        out[tag] = a.directSumWith(b)
    return out


def directSumListsOfSparse(list1, list2):
    return [directSumSparse(a, b) for a, b in zip(list1, list2)]


__all__ = ["Identity", "Complete", "OneSiteOperator", "TwoSiteOperator", "TwoSiteOperatorCompressed",
           "LEFT", "RIGHT", "CENTER", "addStandardCompleteAndIdentityTerms", "contractSparseTensors",
           "directSumListsOfSparse", "directSumSparse", "formSparseContractor", "getInformationFromOperatorCenter",
           "makeSimpleSparseOperator", "makeSparseOperator", "mapOverSparseData", "stripAllButIdentityFrom"]
