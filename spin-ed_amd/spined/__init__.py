"""spined -- MI355X-native exact diagonalisation behind SpinED's ls_* ABI."""
from .internal import *  # noqa: F401,F403
