"""``python -m projecticy_amd.iceflow3d``: the 3D dam-break harness
(iceflow3d/__main__.py)."""

from .harness import main

if __name__ == "__main__":
    main()
