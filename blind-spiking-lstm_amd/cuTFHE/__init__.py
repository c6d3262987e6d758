"""Drop-in for the reference package cuTFHE (GPU engine over the HIP C ABI)."""
