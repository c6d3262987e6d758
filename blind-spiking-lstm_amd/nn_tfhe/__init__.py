"""Homomorphic NN layers mirroring the reference package nn_tfhe."""
