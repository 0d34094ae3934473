"""``python -m pyrometheus_amd``: the code-generation command line
(``pyrometheus/__main__.py``)."""

import sys

from .cli import main

if __name__ == "__main__":
    sys.exit(main())
