"""Contraction compiler, Multiplier, index helpers.  Re-implements the parts of
carcassonne/utils.py on the hot path: Join (45-102), Multiplier (180-207),
prepend decorators (211-227), formDataContractor (562-712), relaxOver (805-878,
see carcassonne_amd.solver), unitize (879-882) and O/L/R (886-888).

formDataContractor precomputes a contraction plan instead of generating source
text: joins are contracted in the order given; after each pairwise contraction
the remaining joins are re-pointed at the new tensor (swapping so the lower
tensor number is on the left) and joins that now connect the same pair are
merged into the earliest of them; leftover tensors are combined by outer
products in ascending order; the result is joined by `final_groups` (or reduced
to a scalar).  These rules fix the leg order of every result, as in the
reference.
"""
from functools import partial

from .errors import DimensionMismatchError, InvariantViolatedError, RelaxFailed, UnexpectedTensorRankError  # noqa: F401


# ---------------------------------------------------------------- index helpers
def O(i): return (i + 2) % 4
def L(i): return (i + 1) % 4
def R(i): return (i - 1) % 4


def dropAt(iterable, index):
    values = (x for i, x in enumerate(iterable) if i != index)
    try:
        return type(iterable)(values)
    except TypeError:
        return tuple(values)


def computeNewDimension(old_dimension, by=None, to=None):
    """utils.py:363-374."""
    if by is None and to is None:
        raise ValueError("Either 'by' or 'to' must not be None.")
    if by is not None and to is not None:
        raise ValueError("Both 'by' ({}) and 'to' ({}) cannot be None.".format(by, to))
    new_dimension = old_dimension + by if by is not None else to
    assert new_dimension >= old_dimension
    return new_dimension


# ---------------------------------------------------------------- Join
class Join:
    def __init__(self, left_tensor_number, left_tensor_indices, right_tensor_number, right_tensor_indices):
        if not hasattr(left_tensor_indices, "__len__") or isinstance(left_tensor_indices, str):
            left_tensor_indices = [left_tensor_indices]
        if not hasattr(right_tensor_indices, "__len__") or isinstance(right_tensor_indices, str):
            right_tensor_indices = [right_tensor_indices]
        left_tensor_indices, right_tensor_indices = list(left_tensor_indices), list(right_tensor_indices)
        if len(left_tensor_indices) != len(right_tensor_indices):
            raise ValueError("number of left indices does not match number of right indices (len({}) != len({}))".format(
                left_tensor_indices, right_tensor_indices))
        self.left_tensor_number, self.left_tensor_indices = left_tensor_number, left_tensor_indices
        self.right_tensor_number, self.right_tensor_indices = right_tensor_number, right_tensor_indices
        self._order()

    def __repr__(self):
        return "Join({},{},{},{})".format(self.left_tensor_number, self.left_tensor_indices,
                                          self.right_tensor_number, self.right_tensor_indices)

    def _order(self):
        if self.left_tensor_number == self.right_tensor_number:
            raise ValueError("attempted to create a self-join for tensor {} (indices = {},{})".format(
                self.left_tensor_number, self.left_tensor_indices, self.right_tensor_indices))
        if self.right_tensor_number < self.left_tensor_number:
            self.left_tensor_number, self.right_tensor_number = self.right_tensor_number, self.left_tensor_number
            self.left_tensor_indices, self.right_tensor_indices = self.right_tensor_indices, self.left_tensor_indices

    checkOrderAndSwap = _order

    def mergeWith(self, other):
        assert (self.left_tensor_number, self.right_tensor_number) == (other.left_tensor_number, other.right_tensor_number)
        self.left_tensor_indices = self.left_tensor_indices + other.left_tensor_indices
        self.right_tensor_indices = self.right_tensor_indices + other.right_tensor_indices

    def update(self, old, new, index_map):
        if self.left_tensor_number == old:
            self.left_tensor_number = new
            self.left_tensor_indices = [index_map[i] for i in self.left_tensor_indices]
            self._order()
        elif self.right_tensor_number == old:
            self.right_tensor_number = new
            self.right_tensor_indices = [index_map[i] for i in self.right_tensor_indices]
            self._order()


def _post_map(rank, contracted, offset=0):
    contracted = set(contracted)
    m, k = {}, 0
    for i in range(rank):
        if i not in contracted:
            m[i] = k + offset
            k += 1
    return m


def _gapless(indices, message):
    if not indices:
        return 0
    n = max(indices) + 1
    missing = set(range(n)) - set(indices)
    if missing:
        raise ValueError(message + ": " + str(missing))
    return n


class _Contractor:
    """A compiled contraction plan; callable on any Data implementation."""

    def __init__(self, ranks, checks, steps, n_inputs, final, outer, groups):
        self.ranks, self.checks, self.steps = ranks, checks, steps
        self.n_inputs, self.final, self.outer, self.groups = n_inputs, final, outer, groups

    def __call__(self, *tensors):
        if len(tensors) != self.n_inputs:
            raise TypeError("contractor takes {} tensors ({} given)".format(self.n_inputs, len(tensors)))
        for k, r in enumerate(self.ranks):
            if tensors[k].ndim != r:
                raise UnexpectedTensorRankError(k, r, tensors[k].ndim)
        for (l, li, r, ri) in self.checks:
            if tensors[l].shape[li] != tensors[r].shape[ri]:
                raise DimensionMismatchError(l, li, tensors[l].shape[li], r, ri, tensors[r].shape[ri])
        slots = list(tensors) + [None] * len(self.steps)
        for k, (l, r, li, ri) in enumerate(self.steps):
            slots[self.n_inputs + k] = slots[l].contractWith(slots[r], li, ri)
            slots[l] = slots[r] = None
        result = slots[self.final]
        for t in self.outer:
            result = result.contractWith(slots[t], [], [])
        if self.groups is None:
            return result.extractScalar()
        if all(len(g) == 1 for g in self.groups) and hasattr(result, "permuted_lazily"):
            # a pure leg permutation of the last product: left to the consumer (stage-3 packing reads it
            # in place through strides; anything else materialises it on first use)
            return result.permuted_lazily([g[0] for g in self.groups])
        return result.join(*self.groups)


def formDataContractor(joins, final_groups, tensor_ranks=None):
    """utils.py:562-712 (see module docstring for the plan rules)."""
    joins = [Join(j.left_tensor_number, list(j.left_tensor_indices), j.right_tensor_number, list(j.right_tensor_indices))
             for j in joins]
    final_groups = [list(g) for g in final_groups]
    observed, pairs = {}, set()

    def observe(t, *indices):
        if t < 0:
            raise ValueError("tensor numbers must be non-negative, not {}".format(t))
        s = observed.setdefault(t, set())
        for i in indices:
            if i in s:
                raise ValueError("index {} of tensor {} appears twice".format(i, t))
            s.add(i)
    for j in joins:
        observe(j.left_tensor_number, *j.left_tensor_indices)
        observe(j.right_tensor_number, *j.right_tensor_indices)
        key = (j.left_tensor_number, j.right_tensor_number)
        if key in pairs:
            raise ValueError("two joins appear between tensors {} and {} (recall that specified joins can contain "
                             "multiple indices)".format(*key))
        pairs.add(key)
    for g in final_groups:
        for (t, i) in g:
            observe(t, i)
    n = _gapless(list(observed.keys()), "the following tensor numbers were expected but not observed")
    if n == 0:
        raise ValueError("no tensors have been specified to be contracted")
    observed_ranks = [_gapless(list(observed[t]), "the following indices of tensor {} were expected but not observed".format(t))
                      for t in range(n)]
    if tensor_ranks is None:
        tensor_ranks = observed_ranks
    elif list(tensor_ranks) != observed_ranks:
        raise ValueError("the ranks of the arguments were specified to be {}, but inferred to be {}".format(
            tensor_ranks, observed_ranks))
    ranks = list(tensor_ranks)
    checks = [(j.left_tensor_number, li, j.right_tensor_number, ri)
              for j in joins for (li, ri) in zip(j.left_tensor_indices, j.right_tensor_indices)]
    steps = []
    active = set(range(n))
    nxt = n
    pending = list(joins)
    while pending:
        j = pending.pop(0)
        l, li, r, ri = j.left_tensor_number, j.left_tensor_indices, j.right_tensor_number, j.right_tensor_indices
        active -= {l, r}
        active.add(nxt)
        ranks.append(ranks[l] + ranks[r] - 2 * len(li))
        steps.append((l, r, list(li), list(ri)))
        lmap = _post_map(ranks[l], li)
        rmap = _post_map(ranks[r], ri, len(lmap))
        for p in pending:
            p.update(l, nxt, lmap)
            p.update(r, nxt, rmap)
        merged = []
        for p in pending:                      # merge later joins into the earliest one on the same pair
            for q in merged:
                if (q.left_tensor_number, q.right_tensor_number) == (p.left_tensor_number, p.right_tensor_number):
                    q.mergeWith(p)
                    break
            else:
                merged.append(p)
        pending = merged
        for g in final_groups:
            for k, (t, i) in enumerate(g):
                if t == l:
                    g[k] = (nxt, lmap[i])
                elif t == r:
                    g[k] = (nxt, rmap[i])
        nxt += 1
    active = sorted(active)
    groups = None
    if final_groups:
        offset, imap = 0, {}
        for t in active:
            for i in range(ranks[t]):
                imap[(t, i)] = i + offset
            offset += ranks[t]
        groups = [[imap[x] for x in g] for g in final_groups]
    return _Contractor(list(tensor_ranks), checks, steps, n, active[0], active[1:], groups)


# ---------------------------------------------------------------- Multiplier + decorators
class Multiplier:
    """utils.py:180-207."""

    def __init__(self, shape, multiply, cost_of_multiply, formMatrix, cost_of_formMatrix):
        self.shape = shape
        self.multiply = multiply
        self.cost_of_multiply = cost_of_multiply
        self.formMatrix = formMatrix
        self.cost_of_formMatrix = cost_of_formMatrix

    def __call__(self, vector):
        return self.multiply(vector)

    @classmethod
    def fromMatrix(cls, matrix):
        m, n = matrix.shape
        return Multiplier(matrix.shape, lambda v: matrix.matvecWith(v), m * n, lambda: matrix, 0)

    def isCheaperToFormMatrix(self, estimated_number_of_applications):
        return estimated_number_of_applications * self.cost_of_multiply > \
            self.cost_of_formMatrix + estimated_number_of_applications * self.shape[0] * self.shape[1]


class prepend:
    def __init__(self, *args, **keywords):
        self.args, self.keywords = args, keywords

    def __call__(self, f):
        return partial(f, *self.args, **self.keywords)


def prependDataContractor(*args, **keywords):
    return prepend(formDataContractor(*args, **keywords))


def multiplyBySingleSiteOperator(state, operator):
    return state.absorbMatrixAt(4, operator)


def unitize(matrix):
    return matrix.unitize()


def relaxOver(*args, **kwargs):
    """utils.py:805-878 on device vectors (see carcassonne_amd.solver.relax_over)."""
    from .solver import relax_over
    return relax_over(*args, **kwargs)


def computeCompressor(old_dimension, new_dimension, multiplier, dtype=None, normalize=False):
    """utils.py:268-303: the eigenvectors of the new_dimension largest eigenvalues of the (Hermitian
    positive semi-definite) Gram matrix behind `multiplier`, as the rows of the compressor
    (new_dimension x old_dimension), and the inverse compressor's conjugate (the same rows, or with
    normalize the rows scaled by sqrt(eval) and 1/sqrt(eval)).  The matrix is formed and decomposed on
    the device (one-sided Jacobi SVD); the reference forms it when new >= old/2 and otherwise runs
    ARPACK on the same eigenpairs."""
    if new_dimension < 0:
        raise ValueError("New dimension ({}) must be non-negative.".format(new_dimension))
    if new_dimension > old_dimension:
        raise ValueError("New dimension ({}) must be less than or equal to the old dimension ({})."
                         .format(new_dimension, old_dimension))
    matrix = multiplier.formMatrix()
    if new_dimension == 0:
        z = type(matrix).newZeros((0, old_dimension))
        return z, z
    if tuple(matrix.shape) != (old_dimension,) * 2:
        raise ValueError("Multiplier matrix has shape {} but the old dimension is {}.".format(matrix.shape, old_dimension))
    # M = V S V^T for a PSD matrix: the rows of V^T are eigenvectors, largest first.  (The left factor
    # is not used: one-sided Jacobi leaves the columns of U for zero singular values undefined, while
    # V, the product of the rotations, is orthogonal whatever the rank.)
    _, S, Vt = matrix.svd(False)
    evals = S.toArray()[:new_dimension]
    if abs(evals[0]) < 1e-15:
        raise ValueError("Input is filled with near-zero elements.")
    evecs = Vt[:new_dimension, :]
    if normalize:
        r = type(matrix).fromArray(np.sqrt(evals).reshape(new_dimension, 1) * np.ones((1, old_dimension)))
        return evecs * r, evecs / r
    return evecs, evecs
