"""srdatalog_hip: MI355X-native (gfx950 HIP) execution backend for srdatalog programs."""
