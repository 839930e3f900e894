"""moldy_amd: MI355X (gfx950) HIP implementation of MOLDY's Fe force hot path."""
