"""The reference's import path, resolved to the MI355X implementation.

Code written against gcross/Carcassonne keeps its imports unchanged:

    from carcassonne.system import System                       (carcassonne/system/__init__.py:2)
    from carcassonne.tensors._2d.sparse import formExpectationAndNormalizationMultipliers
    from carcassonne.utils import relaxOver, Multiplier
    from carcassonne.compression import computeProductCompressor
    from carcassonne.data import NDArrayData
    from carcassonne.policies import ConstantStateCompressionPolicy
    from carcassonne.sparse import Identity, makeSimpleSparseOperator

Each name below is the module object of carcassonne_amd itself (registered in sys.modules under the
reference's dotted name), so `carcassonne.system._2d is carcassonne_amd.system._2d` and every class,
function and module-level state is shared: the numerical work runs in the HIP library
(carcassonne_amd/libcarcassonne_hip.so) exactly as through carcassonne_amd.  `NDArrayData` is the
device-resident data class (HipData).  The reference's 1D modules (system/_1d.py, system/_1d2d.py,
tensors/_1d.py) are not part of the 2D sweep this package implements and are not provided.
"""
import importlib
import sys

# reference module name -> implementing module
_MODULES = (
    ("data", "carcassonne_amd.data"),
    ("data.cost_tracker", "carcassonne_amd.data.cost_tracker"),
    ("utils", "carcassonne_amd.utils"),
    ("sparse", "carcassonne_amd.sparse"),
    ("compression", "carcassonne_amd.compression"),
    ("policies", "carcassonne_amd.policies"),
    ("tensors", "carcassonne_amd.tensors"),
    ("tensors._2d", "carcassonne_amd.tensors._2d"),
    ("tensors._2d.dense", "carcassonne_amd.tensors._2d.dense"),
    ("tensors._2d.sparse", "carcassonne_amd.tensors._2d.sparse"),
    ("system", "carcassonne_amd.system"),
    ("system.base", "carcassonne_amd.system.base"),
    ("system._2d", "carcassonne_amd.system._2d"),
)

for _name, _target in _MODULES:
    _module = importlib.import_module(_target)
    sys.modules[__name__ + "." + _name] = _module
    _parent, _, _leaf = _name.rpartition(".")
    setattr(sys.modules[__name__ + "." + _parent] if _parent else sys.modules[__name__], _leaf, _module)

del _name, _target, _module, _parent, _leaf
