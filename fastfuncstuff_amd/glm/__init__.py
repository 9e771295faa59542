"""GLM engines (ARMA(1,1) REML)."""
