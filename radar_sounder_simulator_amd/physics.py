"""Host-side scalar physics used to set up kernel calls.

``fresnel_normal`` restates ``physics.py:16-21`` (normal-incidence field
reflection coefficient between two relative permittivities).
"""

import numpy as np

C = 299792458.0


def fresnel_normal(eps1, eps2):
    """(n1 - n2)/(n1 + n2) with n = sqrt(eps); air -> ice (3.15) = -0.2792."""
    n1, n2 = np.sqrt(eps1), np.sqrt(eps2)
    return (n1 - n2) / (n1 + n2)
