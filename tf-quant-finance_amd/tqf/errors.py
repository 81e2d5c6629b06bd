"""Exception types matching the reference's error conventions."""


class InvalidArgumentError(ValueError):
  """Stands in for tf.errors.InvalidArgumentError (validate_args asserts)."""
