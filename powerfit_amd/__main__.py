"""``python -m powerfit_amd ...`` runs the powerfit-compatible command line."""
from .cli import main

main()
