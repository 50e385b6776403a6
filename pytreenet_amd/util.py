"""Exception types of the reference (``util/ttn_exceptions.py``)."""


class NotCompatibleException(Exception):
    """Raised when two objects are not compatible (shapes, trees, ids)."""


class NoConnectionException(Exception):
    """Raised when two nodes are not connected."""
