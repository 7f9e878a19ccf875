"""``python -m gpusw_amd <command>``: the ``gpusw`` command line (see :mod:`gpusw_amd.cli`)."""
from .cli import main

raise SystemExit(main())
