"""Host-side helpers around the kernel: seeded scenarios and timers."""
