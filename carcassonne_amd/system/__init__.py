from ._2d import System  # noqa: F401
