"""Callers of the REML search outside ``fit_glm_arma11``."""
