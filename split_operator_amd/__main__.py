"""`python -m split_operator_amd` builds the library in-tree."""
from .build import build

build(verbose=True)
