// Fused contraction and scatter (not built yet).
