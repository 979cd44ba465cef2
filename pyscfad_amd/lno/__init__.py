"""LNO (T) with an occupied projector on MI355X behind pyscfad's ``pyscfad.lno.ccsd_t`` interface."""
from . import ccsd_t  # noqa: F401
