"""CCSD(T) perturbative triples on MI355X behind pyscfad's ``pyscfad.cc.ccsd_t`` interface."""
from . import ccsd_t  # noqa: F401
