"""MI355X-native box fast Gauss transform (drop-in for the reference ``boxfgt``)."""
