"""Host-side checks of the autograd binding (no GPU): render_views rejects
view sets whose viewports differ (the outputs are one stacked tensor and the
library's device readback and adjoint take one size per view)."""
import numpy as np
import pytest

pytest.importorskip("torch")

from micropolis_amd.autograd import check_views  # noqa: E402
from micropolis_amd.renderer import MicropolisError, Projection  # noqa: E402


def test_views_of_one_size_pass():
    mv = np.eye(4, dtype=np.float32)
    check_views([(Projection(60, 0.01, (96, 64)), mv), (Projection(45, 0.1, (96, 64)), mv)])


def test_mixed_viewports_are_rejected():
    mv = np.eye(4, dtype=np.float32)
    with pytest.raises(MicropolisError, match="one viewport"):
        check_views([(Projection(60, 0.01, (96, 64)), mv), (Projection(60, 0.01, (128, 64)), mv)])
    with pytest.raises(MicropolisError):
        check_views([])
