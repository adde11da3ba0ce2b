"""MI355X (gfx950) ray-tracing hot path of RAPTor: C-ABI library + Python harness."""
