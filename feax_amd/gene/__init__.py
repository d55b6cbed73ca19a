"""Responses for optimisation on the HIP path (feax.gene API subset)."""

from .responses import compliance_and_sensitivity, create_compliance_fn

__all__ = ["create_compliance_fn", "compliance_and_sensitivity"]
