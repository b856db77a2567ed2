"""delfem_amd: MI355X-native implementation of DelFEM's linear-solve hot path."""
