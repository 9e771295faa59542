"""GLM engines (ARMA(1,1) REML and the OLS baseline), with the names the
reference's ``fastfuncstuff.glm`` exports for this path (glm/__init__.py:3-17)."""

from .arma import (  # noqa: F401
    ARMA11Results,
    batch_reml_grid_search,
    build_arma11_covariance,
    build_ljung_box_tau,
    compare_ols_vs_arma11,
    compute_arma_lambda,
    compute_ljung_box_statistic,
    fit_glm_arma11,
    ljung_box_max_lag,
    load_arma_params,
    prewhiten_with_arma11,
    reml_grid_search,
    save_arma_rvar,
)
from .core import GLMResults, fit_glm  # noqa: F401
