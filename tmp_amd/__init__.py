"""MI355X-native UnstructSolver residual-and-update path (gfx950 HIP behind a C ABI)."""
