"""``python -m projecticy_amd.iceflow2d``: the falling-ice harness
(iceflow2d/__main__.py)."""

from .harness import main, parse_args

__all__ = ["main", "parse_args"]

if __name__ == "__main__":
    main()
