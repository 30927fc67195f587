"""Small dense linear algebra of the sweep (oracle).

* absorb_matrix_at  -- NDArrayData.absorbMatrixAt (carcassonne/data/__init__.py:148-150)
* normalize_axis    -- NDArrayData.normalizeAxis (data/__init__.py:263-301)
* unitize           -- polar factor U V^H (carcassonne/utils.py:879-882)
* new_enlargener    -- NDArrayData.newEnlargener (data/__init__.py:43-50)
* product_compressor-- computeProductCompressor (carcassonne/compression.py:11-46)
* compute_compressor -- computeCompressor (carcassonne/utils.py:268-303): the leading eigenvectors
                       of a Gram matrix (operator-channel compression)

Pinned by: normalize_axis -- isometry and inverse identities
(reference test_data.py:98-133 restated in tests/test_oracle_linalg.py);
product_compressor -- exact reconstruction when the target rank suffices
(reference test_compression.py:9-30, test_system.py:40-81).  The raw compressor
after four non-converged ALS sweeps is "parity unpinned" (the reference starts
from an unseeded random isometry and uses SciPy GMRES defaults).  compute_compressor --
eigenvalue identities of its output and the operator-compression exactness of
tests/test_oracle_system.py (reference test_system.py:82-177).
"""
import numpy as np
from scipy.linalg import svd


def join(a, *groups):
    """NDArrayData.join (data/__init__.py:247-256): transpose by the flattened groups,
    then merge each group into one axis."""
    groups = [[g] if isinstance(g, (int, np.integer)) else list(g) for g in groups]
    t = a.transpose([i for g in groups for i in g])
    shape, k = [], 0
    for g in groups:
        shape.append(int(np.prod(t.shape[k:k + len(g)], dtype=np.int64)))
        k += len(g)
    return t.reshape(shape)


def absorb_matrix_at(a, axis, m):
    """new[..., i, ...] = sum_j m[i, j] a[..., j, ...]."""
    return np.moveaxis(np.tensordot(m, a, (1, axis)), 0, axis)


def unitize(m):
    u, _, vh = svd(m, full_matrices=False)
    return u @ vh


def normalize_axis(a, axis, sqrt_svals=False, dont_recip_under=1e-14):
    """Returns (isometry, normalizer, denormalizer), or (normalizer, denormalizer)
    when sqrt_svals, exactly as data/__init__.py:263-301."""
    if a.shape[axis] == 1:
        n = np.linalg.norm(a)
        if sqrt_svals:
            n = np.sqrt(n)
            return np.array([[1 / n]]), np.array([[n]])
        return a / n, np.array([[1 / n]]), np.array([[n]])
    others = [i for i in range(a.ndim) if i != axis]
    M = join(a, others, axis)
    if M.shape[0] < M.shape[1]:
        raise ValueError("the total number of degrees of freedom in all other axes ({}) are not enough "
                         "to normalize axis ({}) with dimension ({})".format(M.shape[0], axis, M.shape[1]))
    U, S, V = svd(M, full_matrices=False)
    U = U @ V
    S = S.reshape(-1, 1)
    SI = S.copy()
    if dont_recip_under:
        nz = np.abs(SI) > dont_recip_under
        SI[nz] = 1.0 / SI[nz]
    else:
        SI = 1.0 / SI
    if sqrt_svals:
        S, SI = np.sqrt(S), np.sqrt(SI)
        return (V * SI).conj(), V * S
    shape = [a.shape[i] for i in others] + [U.shape[1]]
    iso = np.moveaxis(U.reshape(shape), -1, axis)
    return iso, V.T @ (V * SI).conj(), V.T.conj() @ (V * S)


def normalize_axis_and_denormalize(a, axis_to_norm, axis_to_denorm, other=None):
    """normalizeAxisAndDenormalize (data/__init__.py:302-310)."""
    if other is None:
        other = a
    if a.shape[axis_to_norm] != other.shape[axis_to_denorm]:
        raise ValueError("Normalized axis and denormalized axis have different sizes")
    iso, _, den = normalize_axis(a, axis_to_norm)
    return iso, absorb_matrix_at(other, axis_to_denorm, den)


def new_enlargener(old, new, rng):
    """newEnlargener (data/__init__.py:43-50): identity when equal, else the Q factor of a
    random new x old matrix (and its conjugate)."""
    if new == old:
        e = np.identity(new)
        return e, e
    q, _ = np.linalg.qr(rng.standard_normal((new, old)))
    return q, q.conj()


def compressor_matrix(Lt, U, Rt):
    """The generated contractor of compression.py:12-25 evaluated on (L, U*, U^H, U^T, R):
    joins L2<->U*0, U^T1<->R1 (via U^H/U^T), U*1<->U^T0, L3<->R2; result
    ((L0,R3),(L1,U^H0)) -- a matrix (l*r) x (old*new) acting on vec(x[old,new])."""
    Uc = U.conj()
    # A[(l,r),(o,n)] = sum L[l,o,p,q] Uc[p,m] UH[n',?]...  written from the join list:
    #  tensors: 0=L(l,o1,o2,w) 1=U*(o2,m) 2=U^H(n,o3) 3=U^T(m,o4) 4=R(o3,o4,w,r)
    #  joins: L2-U*0, U^H1-R0, U^T1-R1, U*1-U^T0, L3-R2 ; groups ((L0,R3),(L1,U^H0))
    UH = U.conj().T
    UT = U.T
    A = np.einsum("lapw,pm,nb,mc,bcwr->lran", Lt, Uc, UH, UT, Rt, optimize=True)
    l, r, a, n = A.shape
    return A.reshape(l * r, a * n)


def sign_normalize_columns(U):
    """Each column negated when its entry of largest magnitude (the first such row) is negative."""
    U = np.array(U, copy=True)
    for j in range(U.shape[1]):
        i = int(np.argmax(np.abs(U[:, j])))
        if U[i, j].real < 0:
            U[:, j] = -U[:, j]
    return U


def compressor_start(Lt, new_dimension):
    """A deterministic start for the compressor iteration (the reference draws a random isometry,
    compression.py:35): the new_dimension leading left singular vectors of L matricized at its
    inward ket leg, (old) x (l * old * w) -- the ket subspace the left tensor uses most -- each with
    the sign that makes its largest entry positive."""
    old = Lt.shape[1]
    m = np.moveaxis(Lt, 1, 0).reshape(old, -1)
    u, _, _ = np.linalg.svd(m, full_matrices=False)
    u = sign_normalize_columns(u)
    if u.shape[1] >= new_dimension:
        return u[:, :new_dimension]
    return unitize(np.hstack([u, np.identity(old)[:, :new_dimension - u.shape[1]]]))


def min_norm_solution(A, b, cutoff=1e-10):
    """The minimum-norm least-squares x of A x = b from the normal equations: the eigenpairs of A^H A with
    eigenvalues above cutoff * the largest (singular values above 1e-5 of the largest, as far as the
    reference's GMRES, relative tolerance 1e-5, resolves them)."""
    G = A.conj().T @ A
    w, V = np.linalg.eigh(G)
    keep = w > cutoff * max(w[-1], 1e-300)
    if not np.any(keep):
        return np.zeros(A.shape[1], dtype=np.result_type(A, b))
    Vk = V[:, keep]
    return Vk @ ((Vk.conj().T @ (A.conj().T @ b)) / w[keep])


def product_compressor(Lt, Rt, new_dimension, rng, sweeps=4, start="random"):
    """computeProductCompressor (compression.py:26-45).  The normal equations
    A^H A x = A^H b are solved for the minimum-norm least-squares x with singular values below 1e-5
    of the largest dropped (min_norm_solution), where the reference calls GMRES (relative tolerance
    1e-5) on them.  start="random" draws the reference's random isometric start; start="svd" uses
    compressor_start.  Returns U^T (new x old)."""
    if Lt.shape[1] != Lt.shape[2]:
        raise ValueError("left inward dimensions do not match (given " + str(Lt.shape) + ")")
    if Rt.shape[0] != Rt.shape[1]:
        raise ValueError("right inward dimensions do not match (given " + str(Rt.shape) + ")")
    if Lt.shape[1] != Rt.shape[1]:
        raise ValueError("left and right shapes are incompatible")
    old = Lt.shape[1]
    b = np.tensordot(Lt, Rt, ((1, 2, 3), (0, 1, 2))).ravel()
    U = compressor_start(Lt, new_dimension) if start == "svd" else unitize(rng.standard_normal((old, new_dimension)))
    for _ in range(sweeps):
        A = compressor_matrix(Lt, U, Rt)
        x = min_norm_solution(A, b)
        U = unitize(x.reshape(old, new_dimension))
    return U.T


def compute_compressor(old_dimension, new_dimension, matrix, normalize=False):
    """computeCompressor (utils.py:268-303) on an explicit Hermitian PSD Gram matrix: the eigenvectors of
    the new_dimension largest eigenvalues (eigh; the reference switches to ARPACK eigsh when
    new < old/2, which targets the same eigenpairs), as rows of the compressor; with normalize the rows
    are scaled by sqrt(eval) and the inverse by 1/sqrt(eval).  Returns (compressor, inverse_conj),
    both new x old."""
    if new_dimension < 0:
        raise ValueError("New dimension ({}) must be non-negative.".format(new_dimension))
    if new_dimension > old_dimension:
        raise ValueError("New dimension ({}) must be less than or equal to the old dimension ({})."
                         .format(new_dimension, old_dimension))
    if new_dimension == 0:
        z = np.zeros((0, old_dimension))
        return z, z
    from scipy.linalg import eigh
    evals, evecs = eigh(matrix)
    evals, evecs = evals[-new_dimension:], evecs[:, -new_dimension:]
    if abs(evals[-1]) < 1e-15:
        raise ValueError("Input is filled with near-zero elements.")
    evecs = evecs.T
    if normalize:
        r = np.sqrt(evals).reshape(new_dimension, 1)
        return evecs * r, evecs / r
    return evecs, evecs
