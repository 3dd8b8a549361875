"""MI355X-native (gfx950 HIP) platform for HiFlow3's local linear algebra."""
