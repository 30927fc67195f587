"""Exception classes of the reference (carcassonne/utils.py:13-41) and the status-code map."""


class DimensionMismatchError(ValueError):
    """utils.py:13-23."""

    def __init__(self, left_tensor_number, left_index, left_dimension, right_tensor_number, right_index,
                 right_dimension):
        self.left_tensor_number, self.left_index, self.left_dimension = left_tensor_number, left_index, left_dimension
        self.right_tensor_number, self.right_index, self.right_dimension = right_tensor_number, right_index, right_dimension
        ValueError.__init__(self, "tensor {}'s index {} has dimension {}, whereas tensor {}'s index {} has dimension {}".format(
            left_tensor_number, left_index, left_dimension, right_tensor_number, right_index, right_dimension))


class InvariantViolatedError(Exception):
    """utils.py:24."""


class RelaxFailed(Exception):
    """utils.py:25-33."""

    def __init__(self, initial_value, final_value):
        Exception.__init__(self)
        self.initial_value, self.final_value = initial_value, final_value

    def __str__(self):
        return "{} --> {}".format(self.initial_value, self.final_value)

    def __repr__(self):
        return "RelaxFailed({},{})".format(self.initial_value, self.final_value)


class UnexpectedTensorRankError(ValueError):
    """utils.py:34-41."""

    def __init__(self, tensor_number, expected_rank, actual_rank):
        self.tensor_number, self.expected_rank, self.actual_rank = tensor_number, expected_rank, actual_rank
        ValueError.__init__(self, "tensor {} was expected to have rank {} but actually has rank {}".format(
            tensor_number, expected_rank, actual_rank))


class HipError(RuntimeError):
    """A HIP runtime failure reported by the native library."""


class LibraryMissingError(ImportError):
    """The native HIP library is not built or cannot be loaded."""
