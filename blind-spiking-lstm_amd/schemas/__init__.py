"""TFHE scheme layer mirroring the reference package schemas."""
